// Ghost-layer periodicity for sweeps that read their neighbours from ghost layers (faces LBM355_FACE_GHOST):
// the copies of lbmpy's LBMPeriodicityHandling (advanced_streaming/communication.py:21-126), which moves,
// after each kernel, the populations leaving the domain through one face to the ghost layer (or, for
// in-place patterns, the values a kernel wrote into a ghost layer back into the interior) at the opposite
// face.  The host side (lbmpy_amd/advanced_streaming/communication.py) turns lbmpy's source / destination
// slices into boxes of element offsets; one launch here copies up to BOXES_PER_LAUNCH boxes, one box per
// blockIdx.y.  Also the streaming-pattern access tables the sweep kernels are compiled from, for the host.
#include <string>

#include "lbm_common.h"

namespace lbm355 {

constexpr int BOXES_PER_LAUNCH = 48;

// passed by value: 48 * 28 B of boxes plus strides, within the kernel-argument limit
struct BoxBatch {
  int64_t src[BOXES_PER_LAUNCH], dst[BOXES_PER_LAUNCH];
  int32_t ex[BOXES_PER_LAUNCH], ey[BOXES_PER_LAUNCH], ez[BOXES_PER_LAUNCH];
  int64_t sx, sy, sz;
};

// Element e of box k sits at (i, j, l) = (e mod ex, (e / ex) mod ey, e / (ex ey)), offset i sx + j sy + l sz
// from the box's source and destination corners.  Sources and destinations of one launch are disjoint
// (checked by the host for lbmpy's slices), so the copies need no order.
template <class E>
__global__ void __launch_bounds__(256) copy_boxes_kernel(E* pdf, BoxBatch b) {
  const int k = blockIdx.y;
  const int64_t ex = b.ex[k], ey = b.ey[k], n = ex * ey * b.ez[k];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % ex, r = e / ex;
    const int64_t o = i * b.sx + (r % ey) * b.sy + (r / ey) * b.sz;
    pdf[b.dst[k] + o] = pdf[b.src[k] + o];
  }
}

template <class E>
int launch_boxes(void* pdf, const int64_t* boxes, int32_t nboxes, int64_t sx, int64_t sy, int64_t sz,
                 hipStream_t s) {
  for (int first = 0; first < nboxes; first += BOXES_PER_LAUNCH) {
    const int n = nboxes - first < BOXES_PER_LAUNCH ? nboxes - first : BOXES_PER_LAUNCH;
    BoxBatch b = {};
    b.sx = sx;
    b.sy = sy;
    b.sz = sz;
    int64_t most = 0;
    for (int k = 0; k < n; ++k) {
      const int64_t* q = boxes + 5 * (int64_t)(first + k);
      b.src[k] = q[0];
      b.dst[k] = q[1];
      b.ex[k] = (int32_t)q[2];
      b.ey[k] = (int32_t)q[3];
      b.ez[k] = (int32_t)q[4];
      const int64_t m = q[2] * q[3] * q[4];
      if (m > most) most = m;
    }
    int64_t blocks = (most + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(copy_boxes_kernel<E>, dim3((unsigned)blocks, (unsigned)n), dim3(256), 0, s,
                       reinterpret_cast<E*>(pdf), b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return lbm355_set_error(LBM355_ERR_LAUNCH, (std::string("launch failed: ") + hipGetErrorString(e)).c_str());
  }
  return LBM355_OK;
}

template <class S>
int fill_access(int32_t pattern, int32_t* out) {
  for (int i = 0; i < S::Q; ++i) {
    const Access a = pattern_access<S>(pattern, i);
    int32_t* o = out + 8 * i;
    o[0] = a.r[0];
    o[1] = a.r[1];
    o[2] = a.r[2];
    o[3] = a.rslot;
    o[4] = a.w[0];
    o[5] = a.w[1];
    o[6] = a.w[2];
    o[7] = a.wslot;
  }
  return S::Q;
}

}  // namespace lbm355

using namespace lbm355;

int lbm355_copy_boxes(int32_t dtype, void* pdf, int64_t span, int64_t stride_x, int64_t stride_y, int64_t stride_z,
                      const int64_t* boxes, int32_t n_boxes, void* stream) {
  if (n_boxes == 0) return LBM355_OK;
  if (!pdf || !boxes || n_boxes < 0 || span <= 0) return lbm355_set_error(LBM355_ERR_ARG, "bad pointer, box count or span");
  if (stride_x < 0 || stride_y < 0 || stride_z < 0) return lbm355_set_error(LBM355_ERR_ARG, "negative stride");
  for (int32_t k = 0; k < n_boxes; ++k) {
    const int64_t* q = boxes + 5 * (int64_t)k;
    if (q[2] < 1 || q[3] < 1 || q[4] < 1 || q[2] > INT32_MAX || q[3] > INT32_MAX || q[4] > INT32_MAX)
      return lbm355_set_error(LBM355_ERR_ARG, ("box " + std::to_string(k) + " has an empty or oversized extent").c_str());
    const int64_t reach = (q[2] - 1) * stride_x + (q[3] - 1) * stride_y + (q[4] - 1) * stride_z;
    for (int side = 0; side < 2; ++side)
      if (q[side] < 0 || q[side] + reach >= span)
        return lbm355_set_error(LBM355_ERR_ARG, ("box " + std::to_string(k) + " reaches outside the array").c_str());
  }
  hipStream_t s = (hipStream_t)stream;
  if (dtype == LBM355_F64) return launch_boxes<uint64_t>(pdf, boxes, n_boxes, stride_x, stride_y, stride_z, s);
  if (dtype == LBM355_F32) return launch_boxes<uint32_t>(pdf, boxes, n_boxes, stride_x, stride_y, stride_z, s);
  if (dtype == LBM355_F16) return launch_boxes<uint16_t>(pdf, boxes, n_boxes, stride_x, stride_y, stride_z, s);
  return lbm355_set_error(LBM355_ERR_ARG, "bad dtype");
}

int32_t lbm355_pattern_access(int32_t stencil, int32_t pattern, int32_t* out) {
  if (!out) return -lbm355_set_error(LBM355_ERR_ARG, "out is NULL");
  if (pattern < LBM355_PULL || pattern > LBM355_PUSH) return -lbm355_set_error(LBM355_ERR_ARG, "bad pattern");
  if (stencil == LBM355_D2Q9) return fill_access<D2Q9>(pattern, out);
  if (stencil == LBM355_D3Q19) return fill_access<D3Q19>(pattern, out);
  if (stencil == LBM355_D3Q27) return fill_access<D3Q27>(pattern, out);
  return -lbm355_set_error(LBM355_ERR_ARG, "unknown stencil");
}
