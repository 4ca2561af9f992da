// ftcc_energy_noalloc: the grand-potential accumulation of ftcc_energy
// (kelvin/ft_cc_energy.py:7-32) for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops).  Its only
// scratch -- the per-block partial sums of the two reductions and the six
// reduced sums -- lives in a file-scope device array, one per GPU, whose
// address is looked up once per device when a context is created.  The
// caller passes the weights and the output on the device; nothing is
// allocated, copied from the host or synchronised.  Calls that overlap on
// one GPU share the array and race.
#include "ftcc_ctx.h"

#include <mutex>

namespace ftcc_core {

namespace {

// 8192 partial triples (the block cap of the grid-stride reduction
// kernels, so every problem size fits) and six reduced sums: 196,656 bytes
// per device
constexpr int SCRATCH_TRIPLES = 8192;
constexpr int SCRATCH_DOUBLES = 3 * SCRATCH_TRIPLES + 8;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

std::mutex scratch_mutex;
double *scratch_ptr[MAX_DEVICES] = {};

__global__ void energy_out_kernel(const double *red, double beta,
                                  double *energy) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    energy[0] = (red[2] + red[5]) / beta;
}

}  // namespace

// Device address of this GPU's scratch array (looked up on first use per
// device, with that device current).
double *noalloc_scratch(int device) {
  require(device >= 0 && device < MAX_DEVICES, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    require(e == hipSuccess && p != nullptr,
            std::string("hipGetSymbolAddress(scratch): ") +
                hipGetErrorString(e),
            FTCC_EHIP);
    scratch_ptr[device] = (double *)p;
  }
  return scratch_ptr[device];
}

}  // namespace ftcc_core

using namespace ftcc_core;

extern "C" int ftcc_energy_noalloc(ftcc_ctx *c, const double *T1,
                                   const double *T2, const double *g_dev,
                                   double beta, int qterm,
                                   double *energy_dev) {
  return guarded(c, [&] {
    require(c->have_ints, "integrals not set", FTCC_ESTATE);
    require(T1 && T2 && g_dev && energy_dev, "null pointer");
    require(beta > 0.0, "beta must be positive");
    int device = 0;
    FTCC_CHECK_HIP(hipGetDevice(&device));
    double *ptr = noalloc_scratch(device);
    double *partials = ptr, *red = ptr + 3 * SCRATCH_TRIPLES;
    // the reductions use one partial triple per block of their kernels;
    // refuse, before any launch, sizes whose blocks the array cannot hold
    const int64_t VO = (int64_t)c->nv * c->no;
    require(reduction_blocks(VO) <= SCRATCH_TRIPLES &&
                reduction_blocks(VO * VO) <= SCRATCH_TRIPLES,
            "ftcc_energy_noalloc: more partial sums than its scratch holds (" +
                std::to_string(SCRATCH_TRIPLES) + " triples)");
    hipStream_t s = c->stream;
    // every partial triple is written before it is read: nothing to clear
    int nb = energy_t1(T1, g_dev, c->fovT, c->ng, c->nv, c->no, partials, s);
    reduce_partials(partials, nb, red, s);
    nb = energy_t2(T1, T2, g_dev, c->oovvT, c->ng, c->nv, c->no, qterm,
                   partials, s);
    reduce_partials(partials, nb, red + 3, s);
    hipLaunchKernelGGL(energy_out_kernel, dim3(1), dim3(64), 0, s, red, beta,
                       energy_dev);
    FTCC_CHECK_HIP(hipGetLastError());
  });
}
