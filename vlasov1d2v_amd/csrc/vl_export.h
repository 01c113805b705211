// Run-file records on the device: the state as float32 values of f itself in
// the reference layout, which is what RunData.save_data stores
// (main/data_management.py:37-52, dataset "pdf", dtype 'f'), and the inverse
// for resuming a run from a record.
//
// One workgroup row per (x element, x node) plane of the ghosted state
// (blockIdx.y), threads over the plane's (NU+2) n (NV+2) n velocity nodes, so
// the background f0(u, v) of the reduced forms is indexed without a modulo.
// HBM-bound: s + 4 bytes per value (s = stored bytes) plus the f0 plane, which
// stays in L2.
#pragma once
#include "vl_common.h"

namespace vl {

template <int S>
__global__ __launch_bounds__(kBlock) void k_export_f32(const typename Store<S>::E *__restrict__ f,
                                                       const double *__restrict__ f0, float *__restrict__ out,
                                                       int64_t plane, double inv_scale) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= plane) return;
    const int64_t i = (int64_t)blockIdx.y * plane + j;
    if constexpr (Store<S>::kDelta)
        out[i] = (float)(f0[j] + (double)Store<S>::ld(f, i) * inv_scale);
    else
        out[i] = (float)f[i];
}

// FP64: f = in.  Reduced forms: g = (in - f0) * scale rounded to nearest
// (the stored form of the record's f).
template <int S>
__global__ __launch_bounds__(kBlock) void k_import_f32(const float *__restrict__ in, const double *__restrict__ f0,
                                                       typename Store<S>::E *__restrict__ f, int64_t plane,
                                                       double scale) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= plane) return;
    const int64_t i = (int64_t)blockIdx.y * plane + j;
    if constexpr (Store<S>::kDelta)
        Store<S>::st(f, i, (float)(((double)in[i] - f0[j]) * scale));
    else
        f[i] = (double)in[i];
}

}  // namespace vl
