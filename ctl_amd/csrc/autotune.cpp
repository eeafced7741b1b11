// Autotuner for the ray-casting launch shape: measures candidate (pixel tile, unroll) shapes on
// the caller's own workload (or a sample of it) and remembers the fastest per workload class.
#include "autotune.h"

#include <cstdio>
#include <map>
#include <mutex>
#include <sstream>

namespace ctltune {

namespace {
std::mutex g_mu;
std::map<std::string, Choice> g_table;

// The candidate shapes before the stacked forms for narrow modules are added.
std::vector<ctlk::LaunchShape> candidateList(bool staged);

} // namespace

std::vector<ctlk::LaunchShape> candidates(bool staged, uint32_t channels)
{
    std::vector<ctlk::LaunchShape> out = candidateList(staged);
    if (channels == 0) return out;
    auto same = [](const ctlk::LaunchShape& x, const ctlk::LaunchShape& y) {
        return x.wave_w == y.wave_w && x.wave_h == y.wave_h && x.block_waves_x == y.block_waves_x &&
               x.block_waves_y == y.block_waves_y && x.unroll == y.unroll && x.blocks_per_cu == y.blocks_per_cu &&
               x.chunk == y.chunk && x.lds_texels == y.lds_texels;
    };
    std::vector<ctlk::LaunchShape> all;
    for (const auto& s : out) {
        all.push_back(s);
        // stacked staged shapes take unroll 4, so two candidates can stack to the same shape
        const ctlk::LaunchShape t = ctlk::narrowModuleShape(s, channels, staged && s.chunk != 0);
        bool dup = t.block_waves_x == s.block_waves_x;
        for (const auto& u : all) dup = dup || same(u, t);
        if (!dup) all.push_back(t);
    }
    return all;
}

namespace {

std::vector<ctlk::LaunchShape> candidateList(bool staged)
{
    std::vector<ctlk::LaunchShape> out;
    if (staged) {
        const uint32_t tiles[2][2] = { { 16, 4 }, { 8, 8 } };
        for (const auto& t : tiles)
            for (uint32_t chunk : { 64u, 48u })
                for (uint32_t texels : { 3072u, 2560u, 6400u })
                    for (uint32_t u : { 3u, 4u }) {
                        ctlk::LaunchShape s = ctlk::defaultShape(true);
                        s.wave_w = t[0];
                        s.wave_h = t[1];
                        s.chunk = chunk;
                        s.lds_texels = texels;
                        s.unroll = u;
                        out.push_back(s);
                    }
        out.push_back(ctlk::defaultShape(false));
        return out;
    }
    const uint32_t tiles[3][2] = { { 8, 8 }, { 16, 4 }, { 4, 16 } };
    for (const auto& t : tiles)
        for (uint32_t u : { 2u, 4u })
            for (uint32_t occ : { 0u, 4u }) {
                ctlk::LaunchShape s = ctlk::defaultShape(false);
                s.wave_w = t[0];
                s.wave_h = t[1];
                s.block_waves_x = 2;
                s.block_waves_y = 2;
                s.unroll = u;
                s.blocks_per_cu = occ;
                out.push_back(s);
            }
    return out;
}

} // namespace

bool tune(const std::string& key, const std::function<float(const ctlk::LaunchShape&)>& run,
          Choice* best, bool staged, uint32_t channels)
{
    Choice b;
    b.ms = -1.0f;
    for (const auto& s : candidates(staged, channels)) {
        run(s); // first launch of a shape warms its code object and caches
        const float ms = run(s);
        if (ms < 0.0f) continue;
        if (b.ms < 0.0f || ms < b.ms) {
            b.shape = s;
            b.ms = ms;
        }
    }
    if (b.ms < 0.0f) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[key] = b;
    if (best) *best = b;
    return true;
}

bool lookup(const std::string& key, ctlk::LaunchShape* shape)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(key);
    if (it == g_table.end()) return false;
    *shape = it->second.shape;
    return true;
}

std::string key(ctlk::Storage st, ctlk::Layout ly, int interpolate, uint32_t channels, uint32_t rows,
                bool staged)
{
    // detector size class: number of 256-pixel steps of the larger side; narrow modules (no wider
    // than a wavefront tile, whose candidates include the stacked shapes) are a class of their own
    const uint32_t cls = ((channels > rows ? channels : rows) + 255) / 256;
    const bool narrow = channels <= 16;
    std::ostringstream os;
    os << "s" << int(st) << "_l" << int(ly) << "_i" << interpolate << "_d" << cls << (narrow ? "_n" : "")
       << (staged ? "_lds" : "");
    return os.str();
}

std::string exportTable()
{
    std::lock_guard<std::mutex> lk(g_mu);
    std::ostringstream os;
    os << "{";
    bool first = true;
    for (const auto& kv : g_table) {
        if (!first) os << ",";
        first = false;
        const auto& s = kv.second.shape;
        os << "\"" << kv.first << "\":{\"tile_w\":" << s.wave_w << ",\"tile_h\":" << s.wave_h
           << ",\"waves_x\":" << s.block_waves_x << ",\"waves_y\":" << s.block_waves_y
           << ",\"unroll\":" << s.unroll << ",\"blocks_per_cu\":" << s.blocks_per_cu << ",\"lds_chunk\":" << s.chunk
           << ",\"lds_texels\":" << s.lds_texels << ",\"ms\":" << kv.second.ms << "}";
    }
    os << "}";
    return os.str();
}

} // namespace ctltune
