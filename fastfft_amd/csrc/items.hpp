// Per-template descriptors of batched launches, shared by the kernels and the
// host code that fills them (plain pointers and integers, no device code).
#pragma once

namespace fastfft_amd {

// One template of a batched row launch (blockIdx.y selects it).  R2C: `in`
// holds `rows` real rows of pitch in_ld with L non-zero values, `out` the
// complex rows (the launch's om).  C2R: `in` holds the complex rows (the
// launch's im), `out` the real rows (the launch's om) and `keys` the argmax
// keys of the rows, or null.
struct RowBatchItem {
    const void*         in;
    void*               out;
    unsigned long long* keys;
    long                in_ld;
    int                 L;
    int                 rows;
};

// One template of a correlation launch: element n of template column c is
// a[c*lda + n*a_es], element j of output column c is out[c*ldo + j*o_es]
// (the strides are shared by all templates of a launch).
struct XcorrItem {
    const void* a;
    void*       out;
    int         L; // non-zero template values per column
    int         unused;
};

} // namespace fastfft_amd
