// Tuning cache of the batched correlation: a process-wide table of launch
// shapes, loaded from and merged into a small JSON file (see tuning.hpp).
#include "tuning.hpp"

#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "../../include/fastfft/fastfft.h"

namespace fastfft_amd {

namespace {

std::mutex                       g_mutex;
bool                             g_loaded = false;
std::string                      g_source;
std::map<std::string, TuneEntry> g_table;

std::string library_dir( ) {
    Dl_info info;
    if ( dladdr(reinterpret_cast<void*>(&fastfft_version), &info) && info.dli_fname ) {
        std::string p = info.dli_fname;
        const auto  s = p.rfind('/');
        return s == std::string::npos ? std::string(".") : p.substr(0, s);
    }
    return ".";
}

int int_field(const std::string& obj, const char* name) {
    const std::string pat = std::string("\"") + name + "\"";
    const auto        at  = obj.find(pat);
    if ( at == std::string::npos )
        return 0;
    const auto colon = obj.find(':', at + pat.size( ));
    return colon == std::string::npos ? 0 : std::atoi(obj.c_str( ) + colon + 1);
}

double real_field(const std::string& obj, const char* name) {
    const std::string pat = std::string("\"") + name + "\"";
    const auto        at  = obj.find(pat);
    if ( at == std::string::npos )
        return 0.0;
    const auto colon = obj.find(':', at + pat.size( ));
    return colon == std::string::npos ? 0.0 : std::atof(obj.c_str( ) + colon + 1);
}

// Entries of a file in the format tuning_store writes; anything else gives no entries.
std::map<std::string, TuneEntry> parse(const std::string& text) {
    std::map<std::string, TuneEntry> out;
    auto                             pos = text.find("\"entries\"");
    if ( pos == std::string::npos )
        return out;
    pos = text.find('{', pos);
    if ( pos == std::string::npos )
        return out;
    pos++;
    while ( true ) {
        const auto k0 = text.find_first_of("\"}", pos);
        if ( k0 == std::string::npos || text[k0] == '}' )
            break;
        const auto k1 = text.find('"', k0 + 1);
        const auto o0 = k1 == std::string::npos ? k1 : text.find('{', k1);
        const auto o1 = o0 == std::string::npos ? o0 : text.find('}', o0);
        if ( o1 == std::string::npos )
            break;
        const std::string key = text.substr(k0 + 1, k1 - k0 - 1);
        const std::string obj = text.substr(o0, o1 - o0 + 1);
        TuneEntry         e;
        e.chunk = int_field(obj, "chunk");
        e.group = int_field(obj, "group");
        e.p     = int_field(obj, "p");
        e.ms    = real_field(obj, "ms");
        if ( e.chunk > 0 )
            out[key] = e;
        pos = o1 + 1;
    }
    return out;
}

std::string read_file(const std::string& path) {
    std::ifstream f(path);
    if ( ! f )
        return std::string( );
    std::stringstream ss;
    ss << f.rdbuf( );
    return ss.str( );
}

void load_locked( ) {
    if ( g_loaded )
        return;
    g_loaded        = true;
    const char* env = std::getenv("FASTFFT_TUNING_CACHE");
    std::string path = env && env[0] ? std::string(env) : library_dir( ) + "/tuning.json";
    std::string text = read_file(path);
    if ( text.empty( ) )
        return;
    g_table  = parse(text);
    g_source = path;
}

void write_file(const std::string& path, const std::map<std::string, TuneEntry>& table) {
    std::ostringstream os;
    os << "{\"format\": \"fastfft_amd batch tuning v1\", \"entries\": {";
    bool first = true;
    for ( const auto& kv : table ) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "{\"chunk\": %d, \"group\": %d, \"p\": %d, \"ms\": %.5f}", kv.second.chunk,
                      kv.second.group, kv.second.p, kv.second.ms);
        os << (first ? "\n  " : ",\n  ") << '"' << kv.first << "\": " << buf;
        first = false;
    }
    os << "\n}}\n";
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp);
        if ( ! f )
            return;
        f << os.str( );
    }
    std::rename(tmp.c_str( ), path.c_str( ));
}

} // namespace

std::string device_arch( ) {
    int dev = 0;
    if ( hipGetDevice(&dev) != hipSuccess )
        return "unknown";
    hipDeviceProp_t prop;
    if ( hipGetDeviceProperties(&prop, dev) != hipSuccess )
        return "unknown";
    std::string a = prop.gcnArchName;
    const auto  c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

bool tuning_lookup(const std::string& key, TuneEntry& out) {
    std::lock_guard<std::mutex> lock(g_mutex);
    load_locked( );
    const auto f = g_table.find(key);
    if ( f == g_table.end( ) )
        return false;
    out = f->second;
    return true;
}

void tuning_store(const std::string& key, const TuneEntry& e) {
    std::lock_guard<std::mutex> lock(g_mutex);
    load_locked( );
    g_table[key]    = e;
    const char* out = std::getenv("FASTFFT_TUNING_WRITE");
    if ( out && out[0] ) {
        auto merged = parse(read_file(out));
        merged[key] = e;
        write_file(out, merged);
    }
}

std::string tuning_source( ) {
    std::lock_guard<std::mutex> lock(g_mutex);
    load_locked( );
    return g_source;
}

} // namespace fastfft_amd
