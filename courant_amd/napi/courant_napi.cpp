// courant_napi.node: the courant.h C ABI presented to Node as the `exports` of a
// WebAssembly instance, the shape web/src/solver/courant.ts:19-85 reads.
//
//   memory.buffer   one fixed host arena (an external ArrayBuffer, never grown, so
//                   views stay valid); every pointer crossing the boundary is a
//                   byte offset into it, 0 meaning NULL
//   malloc / free   first-fit allocation inside the arena
//   cem_*           one wrapper per courant.h function (courant_napi_gen.inc,
//                   generated from include/courant.h by gen_addon.py)
//
// Pointers the library owns are mirrored into the arena: spectra, plan arrays,
// mode fields and strings are copied on every call that returns them; field and
// state arrays (cem_field, cem_state_array) are writable, so their arena copy is
// written back into the library before every call on that simulation and
// refreshed after it: write-then-step and step-then-read through a held view
// work as on the wasm module.  Simulations and mode solvers are small integer
// handles, as wasm32 pointers are numbers.
#include <node_api.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <map>
#include <tuple>
#include <vector>

#include "courant.h"

namespace {

// ---- arena and allocator ----
unsigned char *g_arena = nullptr;
size_t g_arena_bytes = 0;
std::map<uint32_t, uint32_t> g_free;    // offset -> size, address order
std::map<uint32_t, uint32_t> g_used;    // offset -> size
constexpr uint32_t kAlign = 16;

uint32_t arena_alloc(size_t n) {
    if (n == 0) n = 1;
    const size_t need = (n + kAlign - 1) / kAlign * kAlign;
    for (auto it = g_free.begin(); it != g_free.end(); ++it) {
        if (it->second < need) continue;
        const uint32_t off = it->first, size = it->second;
        g_free.erase(it);
        if (size > need) g_free[off + (uint32_t)need] = size - (uint32_t)need;
        g_used[off] = (uint32_t)need;
        return off;
    }
    return 0;
}

void arena_free(uint32_t off) {
    auto u = g_used.find(off);
    if (u == g_used.end()) return;
    uint32_t start = off, size = u->second;
    g_used.erase(u);
    auto next = g_free.lower_bound(start);
    if (next != g_free.end() && start + size == next->first) {
        size += next->second;
        next = g_free.erase(next);
    }
    if (next != g_free.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == start) {
            start = prev->first;
            size += prev->second;
            g_free.erase(prev);
        }
    }
    g_free[start] = size;
}

void *P(uint32_t off) { return off ? g_arena + off : nullptr; }

// ---- handles ----
std::vector<cem_sim *> g_sims(1, nullptr);
struct ModeH {
    cem_mode2d *m;
    int ny, nz;
};
std::vector<ModeH> g_modes(1, ModeH{nullptr, 0, 0});

uint32_t new_handle_sim(cem_sim *s) {
    if (!s) return 0;
    for (size_t h = 1; h < g_sims.size(); h++)
        if (!g_sims[h]) {
            g_sims[h] = s;
            return (uint32_t)h;
        }
    g_sims.push_back(s);
    return (uint32_t)(g_sims.size() - 1);
}

uint32_t new_handle_mode(cem_mode2d *m, int ny, int nz) {
    if (!m) return 0;
    for (size_t h = 1; h < g_modes.size(); h++)
        if (!g_modes[h].m) {
            g_modes[h] = ModeH{m, ny, nz};
            return (uint32_t)h;
        }
    g_modes.push_back(ModeH{m, ny, nz});
    return (uint32_t)(g_modes.size() - 1);
}

// ---- mirrors of library-owned buffers ----
struct Mirror {
    uint32_t off;
    size_t bytes;
    void *lib;
    bool writable;
};
// (owner, kind, index) -> mirror; owner is the sim or mode pointer, or nullptr
std::map<std::tuple<const void *, int, long>, Mirror> g_mirrors;

uint32_t mirror(const void *owner, int kind, long index, const void *lib, size_t bytes, bool writable) {
    if (!lib) return 0;
    auto key = std::make_tuple(owner, kind, index);
    auto it = g_mirrors.find(key);
    if (it != g_mirrors.end() && (it->second.bytes != bytes || it->second.lib != lib)) {
        arena_free(it->second.off);
        g_mirrors.erase(it);
        it = g_mirrors.end();
    }
    if (it == g_mirrors.end()) {
        const uint32_t off = arena_alloc(bytes);
        if (!off) return 0;
        it = g_mirrors.emplace(key, Mirror{off, bytes, const_cast<void *>(lib), writable}).first;
    }
    std::memcpy(g_arena + it->second.off, lib, bytes);
    return it->second.off;
}

// writable arena copies of a simulation -> library (before a call on it)
void sync_in(const void *owner) {
    if (!owner) return;
    for (auto &m : g_mirrors)
        if (std::get<0>(m.first) == owner && m.second.writable)
            std::memcpy(m.second.lib, g_arena + m.second.off, m.second.bytes);
}

// library -> writable arena copies (after a call on it)
void sync_out(const void *owner) {
    if (!owner) return;
    for (auto &m : g_mirrors)
        if (std::get<0>(m.first) == owner && m.second.writable)
            std::memcpy(g_arena + m.second.off, m.second.lib, m.second.bytes);
}

void drop_mirrors(const void *owner) {
    for (auto it = g_mirrors.begin(); it != g_mirrors.end();) {
        if (std::get<0>(it->first) == owner) {
            arena_free(it->second.off);
            it = g_mirrors.erase(it);
        } else {
            ++it;
        }
    }
}

uint32_t mirror_string(int kind, const char *str) {
    return str ? mirror(nullptr, kind, 0, str, std::strlen(str) + 1, false) : 0;
}

size_t grid_points(const cem_sim *s) {
    int nx = 0, ny = 0, nz = 0;
    cem_dims(s, &nx, &ny, &nz);
    return (size_t)(nx + 1) * (ny + 1) * (nz + 1);
}

// plan array lengths (courant.h: ca, cb per grid point; ide, be, ce, dd on nodes;
// idh, bh, ch, d on cells)
size_t plan_len(const cem_sim *s, int which, int comp) {
    int n[3] = {0, 0, 0};
    cem_dims(s, &n[0], &n[1], &n[2]);
    if (comp < 0 || comp > 2) return 0;
    switch (which) {
    case 0: case 1: return grid_points(s);
    case 2: case 4: case 5: case 9: return (size_t)n[comp] + 1;
    case 3: case 6: case 7: case 8: return (size_t)n[comp];
    default: return 0;
    }
}

// ---- argument and result conversion ----
struct Call {
    napi_env e;
    size_t argc = 24;
    napi_value argv[24];
    Call(napi_env env, napi_callback_info ci) : e(env) { napi_get_cb_info(env, ci, &argc, argv, nullptr, nullptr); }
    double num(size_t k) const {
        double v = 0;
        if (k < argc) napi_get_value_double(e, argv[k], &v);
        return v;
    }
    int I(size_t k) const { return (int)num(k); }
    long L(size_t k) const { return (long)num(k); }
    double D(size_t k) const { return num(k); }
    size_t Z(size_t k) const { return (size_t)num(k); }
    void *Ptr(size_t k) const { return P((uint32_t)num(k)); }
    cem_sim *S(size_t k) const {
        const size_t h = (size_t)num(k);
        return h < g_sims.size() ? g_sims[h] : nullptr;
    }
    cem_mode2d *M(size_t k) const {
        const size_t h = (size_t)num(k);
        return h < g_modes.size() ? g_modes[h].m : nullptr;
    }
    const ModeH *MH(size_t k) const {
        const size_t h = (size_t)num(k);
        return h < g_modes.size() && g_modes[h].m ? &g_modes[h] : nullptr;
    }
    void release_sim(size_t k) const {
        const size_t h = (size_t)num(k);
        if (h < g_sims.size()) g_sims[h] = nullptr;
    }
    void release_mode(size_t k) const {
        const size_t h = (size_t)num(k);
        if (h < g_modes.size()) g_modes[h] = ModeH{nullptr, 0, 0};
    }
};

napi_value ret_num(napi_env e, double v) {
    napi_value r;
    napi_create_double(e, v, &r);
    return r;
}
napi_value ret_void(napi_env e) {
    napi_value r;
    napi_get_undefined(e, &r);
    return r;
}

// ---- malloc / free ----
napi_value js_malloc(napi_env e, napi_callback_info ci) {
    Call a(e, ci);
    return ret_num(e, arena_alloc(a.Z(0)));
}
napi_value js_free(napi_env e, napi_callback_info ci) {
    Call a(e, ci);
    arena_free((uint32_t)a.Z(0));
    return ret_void(e);
}

}  // namespace

#include "courant_napi_gen.inc"

namespace {

napi_value init(napi_env e, napi_value exports) {
    size_t mb = 1024;
    if (const char *v = std::getenv("COURANT_ARENA_MB")) mb = (size_t)std::strtoul(v, nullptr, 10);
    if (mb < 16) mb = 16;
    if (mb > 4095) mb = 4095;   // offsets are 32-bit
    g_arena_bytes = mb << 20;
    g_arena = static_cast<unsigned char *>(std::aligned_alloc(4096, g_arena_bytes));
    if (!g_arena) {
        napi_throw_error(e, nullptr, "courant_napi: cannot allocate the arena");
        return nullptr;
    }
    std::memset(g_arena, 0, 4096);
    g_free[kAlign] = (uint32_t)(g_arena_bytes - kAlign);   // offset 0 stays NULL
    napi_value buf, mem;
    napi_create_external_arraybuffer(e, g_arena, g_arena_bytes, nullptr, nullptr, &buf);
    napi_create_object(e, &mem);
    napi_set_named_property(e, mem, "buffer", buf);
    napi_set_named_property(e, exports, "memory", mem);
    std::vector<napi_property_descriptor> d;
    d.push_back({"malloc", nullptr, js_malloc, nullptr, nullptr, nullptr, napi_enumerable, nullptr});
    d.push_back({"free", nullptr, js_free, nullptr, nullptr, nullptr, napi_enumerable, nullptr});
    for (const auto &f : kExports) d.push_back({f.name, nullptr, f.fn, nullptr, nullptr, nullptr, napi_enumerable, nullptr});
    napi_define_properties(e, exports, d.size(), d.data());
    return exports;
}

}  // namespace

NAPI_MODULE(courant_napi, init)
