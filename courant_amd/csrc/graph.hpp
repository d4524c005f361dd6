// Graphs of kernel launches built node by node.
//
// A chain of kernel nodes added with hipGraphAddKernelNode, each depending on the one
// before, then instantiated: the same graph a stream capture of those launches would
// record, but no stream ever enters capture mode.  A capture in one thread changes what
// the runtime allows other threads to do while it lasts (a stream-ordered allocation on
// another thread's stream has failed with "operation not permitted when stream is
// capturing"), and several host threads here build graphs while others launch,
// allocate and free; built graphs have no such window.
#ifndef COURANT_GRAPH_HPP
#define COURANT_GRAPH_HPP

#include <hip/hip_runtime.h>

namespace cem {

class KernelChain {
  public:
    KernelChain() { ok_ = hipGraphCreate(&g_, 0) == hipSuccess; }
    ~KernelChain() {
        if (g_) (void)hipGraphDestroy(g_);
    }
    KernelChain(const KernelChain &) = delete;
    KernelChain &operator=(const KernelChain &) = delete;

    // append kernel<<<grid, block, lds>>>(args...) after the last node; the node keeps a
    // copy of the argument values
    template <typename... P, typename... A>
    void add(void (*kernel)(P...), dim3 grid, dim3 block, unsigned lds, const A &...args) {
        if (!ok_) return;
        append((const void *)kernel, grid, block, lds, static_cast<const P &>(args)...);
    }

    // instantiate the chain; false if any step failed
    bool instantiate(hipGraphExec_t *exec) {
        if (!ok_) return false;
        return hipGraphInstantiate(exec, g_, nullptr, nullptr, 0) == hipSuccess;
    }

  private:
    template <typename... P>
    void append(const void *kernel, dim3 grid, dim3 block, unsigned lds, const P &...args) {
        void *params[sizeof...(P) > 0 ? sizeof...(P) : 1] = {const_cast<void *>(static_cast<const void *>(&args))...};
        hipKernelNodeParams p{};
        p.func = const_cast<void *>(kernel);
        p.gridDim = grid;
        p.blockDim = block;
        p.sharedMemBytes = lds;
        p.kernelParams = params;
        p.extra = nullptr;
        hipGraphNode_t n = nullptr;
        if (hipGraphAddKernelNode(&n, g_, last_ ? &last_ : nullptr, last_ ? 1 : 0, &p) != hipSuccess) {
            ok_ = false;
            return;
        }
        last_ = n;
    }

    hipGraph_t g_ = nullptr;
    hipGraphNode_t last_ = nullptr;
    bool ok_ = false;
};

}  // namespace cem

#endif
