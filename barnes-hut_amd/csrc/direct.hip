// Direct O(N*M) sums, LDS-tiled FP32 (no MFMA: there is no dense contraction).
//
// One thread owns one query; a block of 256 threads streams the sources
// through LDS in tiles of 256 (x, y, z, m) records, so every source is read
// from HBM once per block and broadcast to 256 queries from LDS.
//
// Pair arithmetic follows src/sbh_ffi.cu:431-501 (forward and query-side VJP)
// and src/sbh_ffi.cu:399-429 (potential brute force).
#include "common.hpp"

namespace sbh {
namespace {

constexpr int kBlock = 256;

struct PairParams {
  float eps2;     // max(eps, 0)^2
  float rc;       // cutoff scale (used when CUT)
  float box;      // periodic box length
  float inv_box;  // 1 / box
  int periodic;
};

template <bool CUT>
__device__ __forceinline__ void pair_force(float qx, float qy, float qz, float4 s,
                                           const PairParams& P, float& fx, float& fy,
                                           float& fz) {
  float rx = min_image_rint(qx - s.x, P.periodic, P.box, P.inv_box);
  float ry = min_image_rint(qy - s.y, P.periodic, P.box, P.inv_box);
  float rz = min_image_rint(qz - s.z, P.periodic, P.box, P.inv_box);
  float r2 = rx * rx + ry * ry + rz * rz;
  float inv_r = rsq_softened(r2 + P.eps2 + 1e-12f);
  float inv_r3 = inv_r * inv_r * inv_r;
  float att = 1.0f;
  if constexpr (CUT) {
    float e, d;
    att = attenuation(r2, P.rc, &e, &d);
  }
  float coeff = -s.w * att * inv_r3;
  fx += coeff * rx;
  fy += coeff * ry;
  fz += coeff * rz;
}

template <bool CUT>
__device__ __forceinline__ void pair_vjp(float qx, float qy, float qz, float gx, float gy,
                                         float gz, float4 s, const PairParams& P,
                                         float& ox, float& oy, float& oz) {
  float rx = min_image_rint(qx - s.x, P.periodic, P.box, P.inv_box);
  float ry = min_image_rint(qy - s.y, P.periodic, P.box, P.inv_box);
  float rz = min_image_rint(qz - s.z, P.periodic, P.box, P.inv_box);
  float r2 = rx * rx + ry * ry + rz * rz;
  float inv_r = rsq_softened(r2 + P.eps2 + 1e-12f);
  float inv_r3 = inv_r * inv_r * inv_r;
  float inv_r5 = inv_r3 * inv_r * inv_r;
  float att = 1.0f, cc = 0.0f;
  if constexpr (CUT) {
    float e, dist;
    att = attenuation(r2, P.rc, &e, &dist);
    cc = inv_r3 * dist * e / (kTwoSqrtPi * P.rc * P.rc * P.rc);
  }
  float rg = rx * gx + ry * gy + rz * gz;
  float linear = att * inv_r3;
  float radial = rg * (3.0f * att * inv_r5 + cc);
  ox += -s.w * (linear * gx - radial * rx);
  oy += -s.w * (linear * gy - radial * ry);
  oz += -s.w * (linear * gz - radial * rz);
}

__device__ __forceinline__ void load_tile(float4* tile, const float* __restrict__ pts,
                                          const float* __restrict__ ms, int n, int base,
                                          int dim) {
  int j = base + static_cast<int>(threadIdx.x);
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (j < n) {
    const float* p = pts + static_cast<int64_t>(j) * dim;
    v.x = p[0];
    v.y = p[1];
    v.z = dim == 3 ? p[2] : 0.0f;
    v.w = ms[j];
  }
  tile[threadIdx.x] = v;
}

// PER: periodicity as a template argument (the non-periodic kernels carry no
// minimum-image arithmetic; the sums are the same bits).
template <bool CUT, bool PER>
__global__ __launch_bounds__(kBlock) void k_direct_force(
    const float* __restrict__ pts, const float* __restrict__ ms, int n,
    const float* __restrict__ qs, int nq, float* __restrict__ out, PairParams P) {
  P.periodic = PER ? 1 : 0;
  __shared__ float4 tile[kBlock];
  int i = blockIdx.x * kBlock + threadIdx.x;
  float qx = 0.f, qy = 0.f, qz = 0.f;
  if (i < nq) {
    qx = qs[3 * static_cast<int64_t>(i) + 0];
    qy = qs[3 * static_cast<int64_t>(i) + 1];
    qz = qs[3 * static_cast<int64_t>(i) + 2];
  }
  float fx = 0.f, fy = 0.f, fz = 0.f;
  for (int base = 0; base < n; base += kBlock) {
    load_tile(tile, pts, ms, n, base, 3);
    __syncthreads();
    // Padding records carry m = 0 and contribute exactly zero.
#pragma unroll 8
    for (int k = 0; k < kBlock; ++k) pair_force<CUT>(qx, qy, qz, tile[k], P, fx, fy, fz);
    __syncthreads();
  }
  if (i < nq) {
    out[3 * static_cast<int64_t>(i) + 0] = fx;
    out[3 * static_cast<int64_t>(i) + 1] = fy;
    out[3 * static_cast<int64_t>(i) + 2] = fz;
  }
}

template <bool CUT, bool PER>
__global__ __launch_bounds__(kBlock) void k_direct_vjp(
    const float* __restrict__ pts, const float* __restrict__ ms, int n,
    const float* __restrict__ qs, const float* __restrict__ cot, int nq,
    float* __restrict__ out, PairParams P) {
  P.periodic = PER ? 1 : 0;
  __shared__ float4 tile[kBlock];
  int i = blockIdx.x * kBlock + threadIdx.x;
  float qx = 0.f, qy = 0.f, qz = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
  if (i < nq) {
    int64_t o = 3 * static_cast<int64_t>(i);
    qx = qs[o];
    qy = qs[o + 1];
    qz = qs[o + 2];
    gx = cot[o];
    gy = cot[o + 1];
    gz = cot[o + 2];
  }
  float ox = 0.f, oy = 0.f, oz = 0.f;
  for (int base = 0; base < n; base += kBlock) {
    load_tile(tile, pts, ms, n, base, 3);
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kBlock; ++k)
      pair_vjp<CUT>(qx, qy, qz, gx, gy, gz, tile[k], P, ox, oy, oz);
    __syncthreads();
  }
  if (i < nq) {
    int64_t o = 3 * static_cast<int64_t>(i);
    out[o] = ox;
    out[o + 1] = oy;
    out[o + 2] = oz;
  }
}

// Full adjoint for self-gravity.  For the pair (k = this thread, j = source):
//   v_kj = -m_j (lin c_k - (d.c_k) K d),  v_jk = -m_k (lin c_j - (d.c_j) K d),
//   grad_x[k] += v_kj - v_jk,   grad_m[k] += att s^3 (c_j . d),
// with d = minimage(x_k - x_j), lin = att s^3, K = 3 att s^5 + cutoff term.
// The j == k pair is skipped: its two halves cancel in exact arithmetic but
// not after FMA contraction, where the residual is of order m eps^-3 ulp.
template <bool CUT, bool PER>
__global__ __launch_bounds__(kBlock) void k_direct_full_vjp(
    const float* __restrict__ pts, const float* __restrict__ ms,
    const float* __restrict__ cot, int n, float* __restrict__ gx_out,
    float* __restrict__ gm_out, PairParams P) {
  P.periodic = PER ? 1 : 0;
  __shared__ float4 tp[kBlock];
  __shared__ float4 tc[kBlock];
  int k = blockIdx.x * kBlock + threadIdx.x;
  float xk = 0.f, yk = 0.f, zk = 0.f, mk = 0.f, cxk = 0.f, cyk = 0.f, czk = 0.f;
  if (k < n) {
    int64_t o = 3 * static_cast<int64_t>(k);
    xk = pts[o];
    yk = pts[o + 1];
    zk = pts[o + 2];
    mk = ms[k];
    cxk = cot[o];
    cyk = cot[o + 1];
    czk = cot[o + 2];
  }
  float ax = 0.f, ay = 0.f, az = 0.f, am = 0.f;
  for (int base = 0; base < n; base += kBlock) {
    int j = base + static_cast<int>(threadIdx.x);
    float4 vp = make_float4(0.f, 0.f, 0.f, 0.f), vc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j < n) {
      int64_t o = 3 * static_cast<int64_t>(j);
      vp = make_float4(pts[o], pts[o + 1], pts[o + 2], ms[j]);
      vc = make_float4(cot[o], cot[o + 1], cot[o + 2], 0.f);
    }
    tp[threadIdx.x] = vp;
    tc[threadIdx.x] = vc;
    __syncthreads();
    int cnt = min(kBlock, n - base);
    for (int t = 0; t < cnt; ++t) {
      if (base + t == k) continue;  // self pair: zero in exact arithmetic
      float4 s = tp[t];
      float4 c = tc[t];
      float dx = min_image_rint(xk - s.x, P.periodic, P.box, P.inv_box);
      float dy = min_image_rint(yk - s.y, P.periodic, P.box, P.inv_box);
      float dz = min_image_rint(zk - s.z, P.periodic, P.box, P.inv_box);
      float r2 = dx * dx + dy * dy + dz * dz;
      float inv_r = rsq_softened(r2 + P.eps2 + 1e-12f);
      float s3 = inv_r * inv_r * inv_r;
      float s5 = s3 * inv_r * inv_r;
      float att = 1.0f, cc = 0.0f;
      if constexpr (CUT) {
        float e, dist;
        att = attenuation(r2, P.rc, &e, &dist);
        cc = s3 * dist * e / (kTwoSqrtPi * P.rc * P.rc * P.rc);
      }
      float lin = att * s3;
      float K = 3.0f * att * s5 + cc;
      float dck = dx * cxk + dy * cyk + dz * czk;
      float dcj = dx * c.x + dy * c.y + dz * c.z;
      float a = dck * K, b = dcj * K;
      // v_kj - v_jk
      ax += -s.w * (lin * cxk - a * dx) + mk * (lin * c.x - b * dx);
      ay += -s.w * (lin * cyk - a * dy) + mk * (lin * c.y - b * dy);
      az += -s.w * (lin * czk - a * dz) + mk * (lin * c.z - b * dz);
      am += lin * dcj;
    }
    __syncthreads();
  }
  if (k < n) {
    int64_t o = 3 * static_cast<int64_t>(k);
    gx_out[o] = ax;
    gx_out[o + 1] = ay;
    gx_out[o + 2] = az;
    gm_out[k] = am;
  }
}

// Potential brute force (src/sbh_ffi.cu:399-429), dim 2 or 3.
__global__ __launch_bounds__(kBlock) void k_bruteforce_gravity(
    const float* __restrict__ pts, const float* __restrict__ ms, int n,
    const float* __restrict__ qs, int nq, int dim, float* __restrict__ out, int periodic,
    float box) {
  __shared__ float4 tile[kBlock];
  int i = blockIdx.x * kBlock + threadIdx.x;
  float qx = 0.f, qy = 0.f, qz = 0.f;
  if (i < nq) {
    const float* q = qs + static_cast<int64_t>(i) * dim;
    qx = q[0];
    qy = q[1];
    qz = dim == 3 ? q[2] : 0.0f;
  }
  float acc = 0.0f;
  for (int base = 0; base < n; base += kBlock) {
    load_tile(tile, pts, ms, n, base, dim);
    __syncthreads();
    int cnt = min(kBlock, n - base);
    for (int k = 0; k < cnt; ++k) {
      float4 s = tile[k];
      float dx = s.x - qx, dy = s.y - qy, dz = s.z - qz;
      if (periodic) {
        dx -= box * rintf(dx / box);
        dy -= box * rintf(dy / box);
        dz -= box * rintf(dz / box);
      }
      float dist = sqrtf(dx * dx + dy * dy + dz * dz);
      acc += -s.w / (dist + 1e-1f);
    }
    __syncthreads();
  }
  if (i < nq) out[i] = acc;
}

// Winding-number brute force (src/sbh_gpu.cu:53-64 with brute_force_gpu.h:18-66):
// per query, the compensated (Kahan) sum of signedSolidAngle<3> over all
// oriented area elements, staged through LDS one tile at a time.
__global__ __launch_bounds__(kBlock) void k_bruteforce_wn(
    const float* __restrict__ pts, const float* __restrict__ nrm,
    const float* __restrict__ area, int n, const float* __restrict__ qs, int nq,
    float* __restrict__ out) {
  __shared__ float4 tp[kBlock];
  __shared__ float4 tn[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  float qx = 0.f, qy = 0.f, qz = 0.f;
  if (i < nq) {
    qx = qs[3 * static_cast<int64_t>(i)];
    qy = qs[3 * static_cast<int64_t>(i) + 1];
    qz = qs[3 * static_cast<int64_t>(i) + 2];
  }
  float total = 0.0f, comp = 0.0f;
  for (int base = 0; base < n; base += kBlock) {
    const int j = base + threadIdx.x;
    if (j < n) {
      const int64_t o = 3 * static_cast<int64_t>(j);
      tp[threadIdx.x] = make_float4(pts[o], pts[o + 1], pts[o + 2], area[j]);
      tn[threadIdx.x] = make_float4(nrm[o], nrm[o + 1], nrm[o + 2], 0.f);
    }
    __syncthreads();
    const int cnt = min(kBlock, n - base);
    for (int k = 0; k < cnt; ++k) {
      const float4 p = tp[k], nv = tn[k];
      const float dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
      const float r = sqrtf(dx * dx + dy * dy + dz * dz);
      const float v = p.w * (nv.x * dx + nv.y * dy + nv.z * dz) * 0.07957747154594767f /
                      (r * r * r + 1e-5f);
      const float y = v - comp;
      const float t = total + y;
      comp = (t - total) - y;
      total = t;
    }
    __syncthreads();
  }
  if (i < nq) out[i] = total;
}

// smoothDistExp brute force (src/sbh_gpu.cu:66-77, 287-322): sum over all
// points of exp(-alpha_j |p_j - q|), Kahan-compensated as
// include/brute_force_gpu.h:18-66, points staged through LDS tiles.
__global__ __launch_bounds__(kBlock) void k_bruteforce_smoothdist(
    const float* __restrict__ pts, const float* __restrict__ alpha, int n,
    const float* __restrict__ qs, int nq, float* __restrict__ out) {
  __shared__ float4 tp[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  float qx = 0.f, qy = 0.f, qz = 0.f;
  if (i < nq) {
    qx = qs[3 * static_cast<int64_t>(i)];
    qy = qs[3 * static_cast<int64_t>(i) + 1];
    qz = qs[3 * static_cast<int64_t>(i) + 2];
  }
  float total = 0.0f, comp = 0.0f;
  for (int base = 0; base < n; base += kBlock) {
    const int j = base + threadIdx.x;
    if (j < n) {
      const int64_t o = 3 * static_cast<int64_t>(j);
      tp[threadIdx.x] = make_float4(pts[o], pts[o + 1], pts[o + 2], alpha[j]);
    }
    __syncthreads();
    const int cnt = min(kBlock, n - base);
    for (int k = 0; k < cnt; ++k) {
      const float4 p = tp[k];
      const float dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
      const float v = expf(-p.w * sqrtf(dx * dx + dy * dy + dz * dz));
      const float y = v - comp;
      const float t = total + y;
      comp = (t - total) - y;
      total = t;
    }
    __syncthreads();
  }
  if (i < nq) out[i] = total;
}

PairParams make_params(float softening, float cutoff, int periodic, float box) {
  PairParams P;
  float eps = fmaxf(softening, 0.0f);
  P.eps2 = eps * eps;
  P.rc = cutoff;
  P.periodic = (periodic != 0 && box > 0.0f) ? 1 : 0;
  P.box = box;
  P.inv_box = box > 0.0f ? 1.0f / box : 0.0f;
  return P;
}

// A finite positive cutoff applies the erfc attenuation; +inf gives exactly 1.
inline bool cutoff_active(float cutoff) { return cutoff > 0.0f && std::isfinite(cutoff); }

int check_direct_common(const SbhBuffer* points, const SbhBuffer* masses,
                        const SbhBuffer* queries, float softening, float cutoff,
                        int threads) {
  if (!buffer_ok(points, SBH_DTYPE_F32, -1) || !buffer_ok(queries, SBH_DTYPE_F32, -1) ||
      !buffer_ok(masses, SBH_DTYPE_F32, -1))
    return fail_invalid("points/masses/queries must be float32 device buffers");
  if (points->rank != 2 || queries->rank != 2)
    return fail_invalid("points/queries must be rank-2");
  if (points->dims[1] != 3 || queries->dims[1] != 3)
    return fail_invalid("points/queries must have dim=3");
  if (elem_count(masses) != points->dims[0]) return fail_invalid("masses length mismatch");
  if (softening < 0.0f) return fail_invalid("softening_scale must be >= 0");
  if (!(cutoff > 0.0f)) return fail_invalid("cutoff_scale must be > 0");
  if (threads <= 0 || threads > 1024) return fail_invalid("threads must be in [1, 1024]");
  if (points->dims[0] > INT32_MAX || queries->dims[0] > INT32_MAX)
    return fail_invalid("too many points");
  return SBH_OK;
}

}  // namespace
}  // namespace sbh

using namespace sbh;

extern "C" int sbh_bruteforce_gravity(void* stream, const SbhBuffer* points,
                                      const SbhBuffer* masses, const SbhBuffer* queries,
                                      SbhBuffer* out, int32_t periodic, float box_length,
                                      int32_t threads) {
  clear_error();
  if (!buffer_ok(points, SBH_DTYPE_F32, -1) || !buffer_ok(queries, SBH_DTYPE_F32, -1) ||
      !buffer_ok(masses, SBH_DTYPE_F32, -1) || !buffer_ok(out, SBH_DTYPE_F32, -1))
    return fail_invalid("points/masses/queries/out must be float32 device buffers");
  if (points->rank != 2 || queries->rank != 2)
    return fail_invalid("points/queries must be rank-2");
  int64_t n = points->dims[0], dim = points->dims[1], nq = queries->dims[0];
  if (queries->dims[1] != dim) return fail_invalid("queries dim mismatch");
  if (elem_count(masses) != n) return fail_invalid("masses length mismatch");
  if (dim != 2 && dim != 3) return fail_invalid("only dim=2 or dim=3 supported");
  if (threads <= 0 || threads > 1024) return fail_invalid("threads must be in [1, 1024]");
  if (elem_count(out) != nq) return fail_invalid("out shape mismatch");
  if (n > INT32_MAX || nq > INT32_MAX) return fail_invalid("too many points");
  if (nq == 0) return SBH_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int per = (periodic != 0 && box_length > 0.0f) ? 1 : 0;
  k_bruteforce_gravity<<<grid_for(nq, kBlock), kBlock, 0, s>>>(
      static_cast<const float*>(points->data), static_cast<const float*>(masses->data),
      static_cast<int>(n), static_cast<const float*>(queries->data), static_cast<int>(nq),
      static_cast<int>(dim), static_cast<float*>(out->data), per, box_length);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

extern "C" int sbh_bruteforce_wn(void* stream, const SbhBuffer* points,
                                 const SbhBuffer* normals, const SbhBuffer* areas,
                                 const SbhBuffer* queries, SbhBuffer* out, int32_t threads) {
  clear_error();
  if (!buffer_ok(points, SBH_DTYPE_F32, 2) || points->dims[1] != 3)
    return fail_invalid("points must have shape (N, 3)");
  const int64_t n = points->dims[0];
  if (!buffer_ok(normals, SBH_DTYPE_F32, 2) || normals->dims[0] != n || normals->dims[1] != 3)
    return fail_invalid("normals must have shape (N, 3)");
  if (!buffer_ok(areas, SBH_DTYPE_F32, 1) || areas->dims[0] != n)
    return fail_invalid("areas must have shape (N,)");
  if (!buffer_ok(queries, SBH_DTYPE_F32, 2) || queries->dims[1] != 3)
    return fail_invalid("queries must have shape (M, 3)");
  const int64_t nq = queries->dims[0];
  if (!buffer_ok(out, SBH_DTYPE_F32, 1) || out->dims[0] != nq)
    return fail_invalid("out must have shape (M,)");
  if (threads <= 0 || threads > 1024) return fail_invalid("threads must be in [1, 1024]");
  if (n > INT32_MAX || nq > INT32_MAX) return fail_invalid("too many points");
  if (nq == 0) return SBH_OK;
  k_bruteforce_wn<<<grid_for(nq, kBlock), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      static_cast<const float*>(points->data), static_cast<const float*>(normals->data),
      static_cast<const float*>(areas->data), static_cast<int>(n),
      static_cast<const float*>(queries->data), static_cast<int>(nq),
      static_cast<float*>(out->data));
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

extern "C" int sbh_bruteforce_smoothdist(void* stream, const SbhBuffer* points,
                                         const SbhBuffer* alphas, const SbhBuffer* queries,
                                         SbhBuffer* out, int32_t threads) {
  clear_error();
  if (!buffer_ok(points, SBH_DTYPE_F32, 2) || points->dims[1] != 3)
    return fail_invalid("points must have shape (N, 3)");
  const int64_t n = points->dims[0];
  if (!buffer_ok(alphas, SBH_DTYPE_F32, 1) || alphas->dims[0] != n)
    return fail_invalid("alphas must have shape (N,)");
  if (!buffer_ok(queries, SBH_DTYPE_F32, 2) || queries->dims[1] != 3)
    return fail_invalid("queries must have shape (M, 3)");
  const int64_t nq = queries->dims[0];
  if (!buffer_ok(out, SBH_DTYPE_F32, 1) || out->dims[0] != nq)
    return fail_invalid("out must have shape (M,)");
  if (threads <= 0 || threads > 1024) return fail_invalid("threads must be in [1, 1024]");
  if (n > INT32_MAX || nq > INT32_MAX) return fail_invalid("too many points");
  if (nq == 0) return SBH_OK;
  k_bruteforce_smoothdist<<<grid_for(nq, kBlock), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      static_cast<const float*>(points->data), static_cast<const float*>(alphas->data),
      static_cast<int>(n), static_cast<const float*>(queries->data), static_cast<int>(nq),
      static_cast<float*>(out->data));
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

extern "C" int sbh_direct_force_soft(void* stream, const SbhBuffer* points,
                                     const SbhBuffer* masses, const SbhBuffer* queries,
                                     SbhBuffer* out, float softening_scale,
                                     float cutoff_scale, int32_t periodic,
                                     float box_length, int32_t threads) {
  clear_error();
  int rc = check_direct_common(points, masses, queries, softening_scale, cutoff_scale,
                               threads);
  if (rc != SBH_OK) return rc;
  int64_t n = points->dims[0], nq = queries->dims[0];
  if (!buffer_ok(out, SBH_DTYPE_F32, 2) || out->dims[0] != nq || out->dims[1] != 3)
    return fail_invalid("out shape mismatch");
  if (nq == 0) return SBH_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PairParams P = make_params(softening_scale, cutoff_scale, periodic, box_length);
  auto* pp = static_cast<const float*>(points->data);
  auto* pm = static_cast<const float*>(masses->data);
  auto* pq = static_cast<const float*>(queries->data);
  auto* po = static_cast<float*>(out->data);
  if (cutoff_active(cutoff_scale))
    (P.periodic ? k_direct_force<true, true> : k_direct_force<true, false>)<<<grid_for(nq, kBlock), kBlock, 0, s>>>(pp, pm, (int)n, pq, (int)nq, po, P);
  else
    (P.periodic ? k_direct_force<false, true> : k_direct_force<false, false>)<<<grid_for(nq, kBlock), kBlock, 0, s>>>(pp, pm, (int)n, pq, (int)nq, po, P);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

extern "C" int sbh_direct_force_soft_vjp(void* stream, const SbhBuffer* points,
                                         const SbhBuffer* masses, const SbhBuffer* queries,
                                         const SbhBuffer* cotangent, SbhBuffer* out,
                                         float softening_scale, float cutoff_scale,
                                         int32_t periodic, float box_length,
                                         int32_t threads) {
  clear_error();
  if (points == nullptr || queries == nullptr || cotangent == nullptr ||
      points->rank != 2 || queries->rank != 2 || cotangent->rank != 2)
    return fail_invalid("points/queries/cotangent must be rank-2");
  if (!buffer_ok(cotangent, SBH_DTYPE_F32, 2))
    return fail_invalid("cotangent must be a float32 device buffer");
  if (points->dims[1] != 3 || queries->dims[1] != 3 || cotangent->dims[1] != 3)
    return fail_invalid("points/queries/cotangent must have dim=3");
  if (cotangent->dims[0] != queries->dims[0]) return fail_invalid("cotangent shape mismatch");
  int rc = check_direct_common(points, masses, queries, softening_scale, cutoff_scale,
                               threads);
  if (rc != SBH_OK) return rc;
  int64_t n = points->dims[0], nq = queries->dims[0];
  if (!buffer_ok(out, SBH_DTYPE_F32, 2) || out->dims[0] != nq || out->dims[1] != 3)
    return fail_invalid("out shape mismatch");
  if (nq == 0) return SBH_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PairParams P = make_params(softening_scale, cutoff_scale, periodic, box_length);
  auto* pp = static_cast<const float*>(points->data);
  auto* pm = static_cast<const float*>(masses->data);
  auto* pq = static_cast<const float*>(queries->data);
  auto* pc = static_cast<const float*>(cotangent->data);
  auto* po = static_cast<float*>(out->data);
  if (cutoff_active(cutoff_scale))
    (P.periodic ? k_direct_vjp<true, true> : k_direct_vjp<true, false>)<<<grid_for(nq, kBlock), kBlock, 0, s>>>(pp, pm, (int)n, pq, pc, (int)nq, po, P);
  else
    (P.periodic ? k_direct_vjp<false, true> : k_direct_vjp<false, false>)<<<grid_for(nq, kBlock), kBlock, 0, s>>>(pp, pm, (int)n, pq, pc, (int)nq, po, P);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

extern "C" int sbh_direct_force_soft_full_vjp(void* stream, const SbhBuffer* points,
                                              const SbhBuffer* masses,
                                              const SbhBuffer* cotangent,
                                              SbhBuffer* grad_points,
                                              SbhBuffer* grad_masses,
                                              float softening_scale, float cutoff_scale,
                                              int32_t periodic, float box_length,
                                              int32_t threads) {
  clear_error();
  int rc = check_direct_common(points, masses, points, softening_scale, cutoff_scale,
                               threads);
  if (rc != SBH_OK) return rc;
  int64_t n = points->dims[0];
  if (!buffer_ok(cotangent, SBH_DTYPE_F32, 2) || cotangent->dims[0] != n ||
      cotangent->dims[1] != 3)
    return fail_invalid("cotangent shape mismatch");
  if (!buffer_ok(grad_points, SBH_DTYPE_F32, 2) || grad_points->dims[0] != n ||
      grad_points->dims[1] != 3)
    return fail_invalid("grad_points shape mismatch");
  if (!buffer_ok(grad_masses, SBH_DTYPE_F32, -1) || elem_count(grad_masses) != n)
    return fail_invalid("grad_masses shape mismatch");
  if (n == 0) return SBH_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PairParams P = make_params(softening_scale, cutoff_scale, periodic, box_length);
  auto* pp = static_cast<const float*>(points->data);
  auto* pm = static_cast<const float*>(masses->data);
  auto* pc = static_cast<const float*>(cotangent->data);
  auto* gx = static_cast<float*>(grad_points->data);
  auto* gm = static_cast<float*>(grad_masses->data);
  if (cutoff_active(cutoff_scale))
    (P.periodic ? k_direct_full_vjp<true, true> : k_direct_full_vjp<true, false>)<<<grid_for(n, kBlock), kBlock, 0, s>>>(pp, pm, pc, (int)n, gx, gm, P);
  else
    (P.periodic ? k_direct_full_vjp<false, true> : k_direct_full_vjp<false, false>)<<<grid_for(n, kBlock), kBlock, 0, s>>>(pp, pm, pc, (int)n, gx, gm, P);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}
