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
// handles, as wasm32 pointers are numbers.  All of it is per Node environment:
// the main thread and every worker thread that loads the addon get their own
// arena and handles, as each would instantiate its own wasm module.
#include <node_api.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <cmath>
#include <initializer_list>
#include <map>
#include <tuple>
#include <vector>

#include "courant.h"

namespace {

constexpr uint32_t kAlign = 16;

// ---- one instance per Node environment (main thread, each worker thread) ----
struct ModeH {
    cem_mode2d *m;
    int ny, nz;
};
struct Mirror {
    uint32_t off;
    size_t bytes;
    void *lib;
    bool writable;
};

struct Instance {
    unsigned char *arena = nullptr;
    size_t arena_bytes = 0;
    std::map<uint32_t, uint32_t> free_blocks;   // offset -> size, address order
    std::map<uint32_t, uint32_t> used;          // offset -> size
    std::vector<cem_sim *> sims{nullptr};
    std::vector<ModeH> modes{ModeH{nullptr, 0, 0}};
    // (owner, kind, index) -> mirror; owner is the sim or mode pointer, or nullptr
    std::map<std::tuple<const void *, int, long>, Mirror> mirrors;

    ~Instance() { std::free(arena); }

    uint32_t alloc(size_t n) {
        if (n == 0) n = 1;
        const size_t need = (n + kAlign - 1) / kAlign * kAlign;
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < need) continue;
            const uint32_t off = it->first, size = it->second;
            free_blocks.erase(it);
            if (size > need) free_blocks[off + (uint32_t)need] = size - (uint32_t)need;
            used[off] = (uint32_t)need;
            return off;
        }
        return 0;
    }

    void release(uint32_t off) {
        auto u = used.find(off);
        if (u == used.end()) return;
        uint32_t start = off, size = u->second;
        used.erase(u);
        auto next = free_blocks.lower_bound(start);
        if (next != free_blocks.end() && start + size == next->first) {
            size += next->second;
            next = free_blocks.erase(next);
        }
        if (next != free_blocks.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                size += prev->second;
                free_blocks.erase(prev);
            }
        }
        free_blocks[start] = size;
    }

    void *ptr(uint32_t off) const { return off ? arena + off : nullptr; }

    uint32_t new_sim(cem_sim *s) {
        if (!s) return 0;
        for (size_t h = 1; h < sims.size(); h++)
            if (!sims[h]) {
                sims[h] = s;
                return (uint32_t)h;
            }
        sims.push_back(s);
        return (uint32_t)(sims.size() - 1);
    }

    uint32_t new_mode(cem_mode2d *m, int ny, int nz) {
        if (!m) return 0;
        for (size_t h = 1; h < modes.size(); h++)
            if (!modes[h].m) {
                modes[h] = ModeH{m, ny, nz};
                return (uint32_t)h;
            }
        modes.push_back(ModeH{m, ny, nz});
        return (uint32_t)(modes.size() - 1);
    }

    uint32_t mirror(const void *owner, int kind, long index, const void *lib, size_t bytes, bool writable) {
        if (!lib) return 0;
        auto key = std::make_tuple(owner, kind, index);
        auto it = mirrors.find(key);
        if (it != mirrors.end() && (it->second.bytes != bytes || it->second.lib != lib)) {
            release(it->second.off);
            mirrors.erase(it);
            it = mirrors.end();
        }
        if (it == mirrors.end()) {
            const uint32_t off = alloc(bytes);
            if (!off) return 0;
            it = mirrors.emplace(key, Mirror{off, bytes, const_cast<void *>(lib), writable}).first;
        }
        std::memcpy(arena + it->second.off, lib, bytes);
        return it->second.off;
    }

    // writable arena copies of a simulation -> library (before a call on it)
    void sync_in(const void *owner) {
        if (!owner) return;
        for (auto &m : mirrors)
            if (std::get<0>(m.first) == owner && m.second.writable)
                std::memcpy(m.second.lib, arena + m.second.off, m.second.bytes);
    }

    // library -> writable arena copies (after a call on it)
    void sync_out(const void *owner) {
        if (!owner) return;
        for (auto &m : mirrors)
            if (std::get<0>(m.first) == owner && m.second.writable)
                std::memcpy(arena + m.second.off, m.second.lib, m.second.bytes);
    }

    void drop_mirrors(const void *owner) {
        for (auto it = mirrors.begin(); it != mirrors.end();) {
            if (std::get<0>(it->first) == owner) {
                release(it->second.off);
                it = mirrors.erase(it);
            } else {
                ++it;
            }
        }
    }

    uint32_t mirror_string(int kind, const char *str) {
        return str ? mirror(nullptr, kind, 0, str, std::strlen(str) + 1, false) : 0;
    }
};

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
    Instance *in = nullptr;
    size_t argc = 24;
    napi_value argv[24];
    Call(napi_env env, napi_callback_info ci) : e(env) {
        napi_get_cb_info(env, ci, &argc, argv, nullptr, nullptr);
        napi_get_instance_data(env, reinterpret_cast<void **>(&in));
    }
    double num(size_t k) const {
        double v = 0;
        if (k < argc) napi_get_value_double(e, argv[k], &v);
        return v;
    }
    int I(size_t k) const { return (int)num(k); }
    long L(size_t k) const { return (long)num(k); }
    double D(size_t k) const { return num(k); }
    size_t Z(size_t k) const { return (size_t)num(k); }
    void *Ptr(size_t k) const { return in->ptr((uint32_t)num(k)); }
    // Every pointer argument must be 0 (NULL) or an offset inside a block of the arena
    // that malloc or a returned array handed out (wasm memory would trap on anything
    // else); otherwise a RangeError is thrown and the library is not called.
    bool ptrs_ok(std::initializer_list<size_t> ks) const {
        for (size_t k : ks) {
            const double v = num(k);
            if (v == 0) continue;
            bool ok = std::isfinite(v) && v > 0 && v == std::floor(v) && v < (double)in->arena_bytes;
            if (ok) {
                const uint32_t off = (uint32_t)v;
                auto it = in->used.upper_bound(off);
                ok = it != in->used.begin() && off < std::prev(it)->first + std::prev(it)->second;
            }
            if (!ok) {
                napi_throw_range_error(e, nullptr, "courant: pointer argument outside the arena's allocated blocks");
                return false;
            }
        }
        return true;
    }
    cem_sim *S(size_t k) const {
        const size_t h = (size_t)num(k);
        return h < in->sims.size() ? in->sims[h] : nullptr;
    }
    cem_mode2d *M(size_t k) const {
        const size_t h = (size_t)num(k);
        return h < in->modes.size() ? in->modes[h].m : nullptr;
    }
    const ModeH *MH(size_t k) const {
        const size_t h = (size_t)num(k);
        return h < in->modes.size() && in->modes[h].m ? &in->modes[h] : nullptr;
    }
    void release_sim(size_t k) const {
        const size_t h = (size_t)num(k);
        if (h < in->sims.size()) in->sims[h] = nullptr;
    }
    void release_mode(size_t k) const {
        const size_t h = (size_t)num(k);
        if (h < in->modes.size()) in->modes[h] = ModeH{nullptr, 0, 0};
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
    return ret_num(e, a.in->alloc(a.Z(0)));
}
napi_value js_free(napi_env e, napi_callback_info ci) {
    Call a(e, ci);
    a.in->release((uint32_t)a.Z(0));
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
    Instance *in = new Instance();
    in->arena_bytes = mb << 20;
    in->arena = static_cast<unsigned char *>(std::aligned_alloc(4096, in->arena_bytes));
    if (!in->arena) {
        delete in;
        napi_throw_error(e, nullptr, "courant_napi: cannot allocate the arena");
        return nullptr;
    }
    std::memset(in->arena, 0, 4096);
    in->free_blocks[kAlign] = (uint32_t)(in->arena_bytes - kAlign);   // offset 0 stays NULL
    // one instance per environment (each worker thread loads its own), freed with it
    napi_set_instance_data(
        e, in, [](napi_env, void *data, void *) { delete static_cast<Instance *>(data); }, nullptr);
    napi_value buf, mem;
    napi_create_external_arraybuffer(e, in->arena, in->arena_bytes, nullptr, nullptr, &buf);
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
