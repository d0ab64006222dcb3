// The discrete adjoint of the RK4 step as a torch::autograd::Function (libtorch
// C++ API) over the C ABI of libdgpar (include/dg_capi.h); no counterpart in the
// reference.
//
//   UT = rk4(U0, ctx, dt, nsteps, ckpt_every)
//
// forward  sets the context's state to U0 and runs dg_adjoint_forward: nsteps
//          RK4 steps of the unlimited FP64 residual (TimeIntegrator.cc:81-275),
//          keeping a checkpoint every ckpt_every steps on the device;
// backward runs dg_adjoint_backward with the incoming gradient dL/dUT as seed:
//          each checkpoint segment's stage states are recomputed, then the fused
//          reverse RK4 stages give dL/dU0.
//
// Tensors cross the boundary as host FP64 tensors in the caller layout [e][f][k]
// (the library keeps its own device buffers; torch's and the library's HIP
// runtimes do not share device pointers).  A context holds one set of
// checkpoints: when another forward ran on the same context before this node's
// backward, the backward re-runs its own forward from the saved U0 first.
#include <torch/extension.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/dg_capi.h"

namespace {

std::mutex g_mu;
std::map<std::uintptr_t, int64_t> g_generation;  // last forward run on each context

void check(dg_ctx* c, dg_status s, const char* what) {
  if (s != DG_OK) throw std::runtime_error(std::string(what) + ": " + dg_last_error(c));
}

torch::Tensor host_f64(const torch::Tensor& t, int64_t n, const char* name) {
  TORCH_CHECK(t.numel() == n, name, " must hold ", n, " values (the context's ndof)");
  return t.detach().to(torch::kCPU, torch::kFloat64).contiguous();
}

int64_t run_forward(dg_ctx* c, const torch::Tensor& U0, double dt, int64_t nsteps, int64_t ckpt) {
  check(c, dg_set_state(c, U0.data_ptr<double>(), (size_t)U0.numel()), "dg_set_state");
  check(c, dg_adjoint_forward(c, dt, (int)nsteps, (int)ckpt, nullptr, nullptr), "dg_adjoint_forward");
  std::lock_guard<std::mutex> lk(g_mu);
  return ++g_generation[(std::uintptr_t)c];
}

struct RK4Adjoint : public torch::autograd::Function<RK4Adjoint> {
  static torch::Tensor forward(torch::autograd::AutogradContext* actx, const torch::Tensor& U0, int64_t handle,
                               double dt, int64_t nsteps, int64_t ckpt) {
    dg_ctx* c = reinterpret_cast<dg_ctx*>(handle);
    TORCH_CHECK(c != nullptr, "null dg context");
    const int64_t n = (int64_t)dg_ndof(c);
    torch::Tensor u0 = host_f64(U0, n, "U0");
    const int64_t gen = run_forward(c, u0, dt, nsteps, ckpt);
    torch::Tensor UT = torch::empty({n}, torch::dtype(torch::kFloat64));
    check(c, dg_get_state(c, UT.data_ptr<double>(), (size_t)n), "dg_get_state");
    actx->save_for_backward({u0});
    actx->saved_data["handle"] = handle;
    actx->saved_data["dt"] = dt;
    actx->saved_data["nsteps"] = nsteps;
    actx->saved_data["ckpt"] = ckpt;
    actx->saved_data["gen"] = gen;
    return UT.to(U0.device()).reshape(U0.sizes());
  }

  static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* actx,
                                               torch::autograd::tensor_list grads) {
    dg_ctx* c = reinterpret_cast<dg_ctx*>(actx->saved_data["handle"].toInt());
    const torch::Tensor u0 = actx->get_saved_variables()[0];
    const int64_t n = u0.numel();
    int64_t gen = actx->saved_data["gen"].toInt();
    {
      std::lock_guard<std::mutex> lk(g_mu);
      if (g_generation[(std::uintptr_t)c] != gen) gen = -1;
    }
    if (gen < 0)
      run_forward(c, u0, actx->saved_data["dt"].toDouble(), actx->saved_data["nsteps"].toInt(),
                  actx->saved_data["ckpt"].toInt());
    torch::Tensor seed = host_f64(grads[0], n, "dL/dUT");
    torch::Tensor g = torch::empty({n}, torch::dtype(torch::kFloat64));
    check(c, dg_adjoint_backward(c, seed.data_ptr<double>(), g.data_ptr<double>()), "dg_adjoint_backward");
    return {g.to(grads[0].device()).reshape(grads[0].sizes()), torch::Tensor(), torch::Tensor(), torch::Tensor(),
            torch::Tensor()};
  }
};

torch::Tensor rk4(const torch::Tensor& U0, int64_t handle, double dt, int64_t nsteps, int64_t ckpt_every) {
  return RK4Adjoint::apply(U0, handle, dt, nsteps, ckpt_every);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "Discrete RK4 adjoint of the DG residual as a torch autograd Function (libdgpar C ABI)";
  m.def("rk4", &rk4, "UT = rk4(U0, ctx_handle, dt, nsteps, ckpt_every); differentiable in U0", pybind11::arg("U0"),
        pybind11::arg("handle"), pybind11::arg("dt"), pybind11::arg("nsteps"), pybind11::arg("ckpt_every"));
}
