// Host-only driver of the planner (fastfft_amd/csrc/planner.cpp) for the
// AddressSanitizer / UndefinedBehaviorSanitizer build in tests/test_asan.py:
// validates and plans every combination of sizes, ranks and windows below,
// builds the three stage lists and their workspace, and checks that every
// stage's line maps stay inside the buffers the plan's sizes describe.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../fastfft_amd/csrc/plan.hpp"

using namespace fastfft_amd;

static long extent(const HostLineMap& m, int nlines, int nelem) {
    if ( nlines <= 0 || nelem <= 0 )
        return 0;
    const long l1 = nlines / (m.n2 > 0 ? m.n2 : 1);
    return (l1 > 0 ? (l1 - 1) * m.s1 : 0) + (m.n2 - 1) * m.s2 + (long)(nelem - 1) * m.es + 1;
}

int main( ) {
    const int sizes[]   = {2, 8, 16, 64, 100, 128, 384, 512, 1024, 4096};
    const int windows[] = {1, 16, 64, 128, 512};
    int       planned = 0, rejected = 0, bad = 0;
    for ( int rank : {2, 3} )
        for ( int in : sizes )
            for ( int n : sizes )
                for ( int m : windows ) {
                    if ( rank == 3 && (n > 512 || in > 128 || m > 128) )
                        continue;
                    for ( int origin : {0, -m / 2} ) {
                        PlanState p;
                        p.input_type = 0;
                        p.image_type = 0;
                        p.rank       = rank;
                        p.fwd_set = p.inv_set = true;
                        const int z  = rank == 3 ? 1 : 0;
                        p.fwd_in     = Dims{in, in, z ? in : 1, 0};
                        p.fwd_out    = Dims{n, n, z ? n : 1, 0};
                        p.inv_in     = Dims{n, n, z ? n : 1, 0};
                        p.inv_out    = Dims{m, m, z ? m : 1, 0};
                        p.origin[0] = p.origin[1] = origin;
                        p.origin[2]               = z ? origin : 0;
                        p.input_memory            = padded_memory_size(p.fwd_in);
                        p.fwd_output_memory       = padded_memory_size(p.fwd_out);
                        (void)padded_memory_size(p.inv_in);
                        p.inv_output_memory = padded_memory_size(p.inv_out);
                        if ( validate(p).code ) {
                            rejected++;
                            continue;
                        }
                        size_t wa = 0, wb = 0;
                        workspace_elems(p, wa, wb);
                        for ( const char* op : {"fwd", "inv", "xcorr"} ) {
                            std::vector<Stage> st;
                            if ( build_stages(p, op, st).code )
                                continue;
                            planned++;
                            (void)describe(p, st);
                            for ( const Stage& s : st ) {
                                const int nw = s.kind == StageKind::fwd_real || s.kind == StageKind::fwd_cpx ? s.nout : s.M;
                                const long ex = extent(s.om, s.nlines, nw);
                                const long cap = s.dst == Buf::ws_a ? (long)wa : s.dst == Buf::ws_b ? (long)wb : -1;
                                if ( cap >= 0 && ex > cap ) {
                                    std::printf("stage %s of %s writes %ld > %ld elements\n", s.ref_kernel.c_str( ), op, ex, cap);
                                    bad++;
                                }
                                if ( s.N < 1 || s.P < 1 || s.Q < 1 || (long)s.P * s.Q != s.N ) {
                                    std::printf("stage %s of %s: N %d != P %d x Q %d\n", s.ref_kernel.c_str( ), op, s.N, s.P, s.Q);
                                    bad++;
                                }
                            }
                        }
                    }
                }
    std::printf("planned %d stage lists, rejected %d plans, %d problems\n", planned, rejected, bad);
    return bad == 0 && planned > 0 ? 0 : 1;
}
