// Tuning cache of the batched correlation (host code only).
//
// The batch's launch shape is chosen per (device architecture, precision,
// element types, transform and frame sizes, largest template, window, batch
// size) by timing candidates on the device (capi.cpp); the choice is kept in
// a process-wide table, filled at first use from a JSON file:
//   FASTFFT_TUNING_CACHE=<path>, else tuning.json beside libfastfft_amd.so.
// With FASTFFT_TUNING_WRITE=<path>, every choice timed in this process is
// merged into that file.  File format (written by tuning_store):
//   {"format": "fastfft_amd batch tuning v1", "entries": {
//     "<key>": {"chunk": C, "group": G, "p": P, "ms": T}, ...}}
#pragma once

#include <string>

namespace fastfft_amd {

struct TuneEntry {
    int    chunk = 0; // templates per launch
    int    group = 0; // templates per template group of the correlation pass
    int    p     = 0; // sub-transform length of the correlation pass (0: the planner's)
    double ms    = 0; // time of one call when the entry was timed
};

// Architecture name of the current device ("gfx950"), or "unknown".
std::string device_arch( );

bool tuning_lookup(const std::string& key, TuneEntry& out);
// Records an entry in the process table (and in FASTFFT_TUNING_WRITE's file when set).
void tuning_store(const std::string& key, const TuneEntry& e);
// Path the table was loaded from ("" when none was found).
std::string tuning_source( );

} // namespace fastfft_amd
