// lj_device.h — device-side building blocks shared by the force kernels.
//
// Layouts: AoS positions/momenta are double[N][4] (x, y, z, pad; 32 B per
// particle, the reference's double4), packed AoS double[N][3] (the
// reference's double3), SoA three double[N] arrays.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ljgpu {

// Number of XCDs on MI355X. Workgroups are dispatched round-robin over them
// (workgroup b runs on XCD b % 8), each XCD with its own L2.
constexpr int kXcds = 8;

// Logical block index for an XCD-aware launch over a grid padded to a
// multiple of kXcds: XCD x receives the contiguous logical range
// [x*per, (x+1)*per), so particles with consecutive indices (which share most
// neighbours in lattice order) are served by the same L2.
__device__ __forceinline__ int xcd_logical_block(int nblocks_padded) {
  const int b = static_cast<int>(blockIdx.x);
  const int per = nblocks_padded / kXcds;
  return (b % kXcds) * per + b / kXcds;
}

// Gathers use 32-bit unsigned byte offsets from a uniform base pointer, so
// the compiler can emit the SGPR-base form of global loads (one VALU shift
// per gather instead of 64-bit address arithmetic). Valid for N < 2^27.
struct PosAos {
  const double* __restrict__ q;
  __device__ __forceinline__ void ld(int j, double& x, double& y, double& z) const {
    const uint32_t off = static_cast<uint32_t>(j) << 5;
    const char* base = reinterpret_cast<const char*>(q);
    const double2 xy = *reinterpret_cast<const double2*>(base + off);
    x = xy.x;
    y = xy.y;
    z = *reinterpret_cast<const double*>(base + off + 16u);
  }
};

// Packed AoS: double[N][3] (x, y, z; 24 B per particle, the reference's
// double3). A row is only 8-byte aligned, so each coordinate is its own
// 8-byte load; 32-bit byte offsets as above (valid for N < 2^32 / 24).
struct PosAos3 {
  const double* __restrict__ q;
  __device__ __forceinline__ void ld(int j, double& x, double& y, double& z) const {
    const uint32_t off = static_cast<uint32_t>(j) * 24u;
    const char* base = reinterpret_cast<const char*>(q);
    x = *reinterpret_cast<const double*>(base + off);
    y = *reinterpret_cast<const double*>(base + off + 8u);
    z = *reinterpret_cast<const double*>(base + off + 16u);
  }
};

struct PosSoa {
  const double* __restrict__ x;
  const double* __restrict__ y;
  const double* __restrict__ z;
  __device__ __forceinline__ void ld(int j, double& a, double& b, double& c) const {
    const uint32_t off = static_cast<uint32_t>(j) << 3;
    a = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(x) + off);
    b = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(y) + off);
    c = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(z) + off);
  }
};

struct MomAos {
  double* __restrict__ p;
  __device__ __forceinline__ void add(int i, double ax, double ay, double az) const {
    double* r = p + 4 * static_cast<int64_t>(i);
    double2 xy = *reinterpret_cast<double2*>(r);
    xy.x += ax;
    xy.y += ay;
    *reinterpret_cast<double2*>(r) = xy;
    r[2] += az;
  }
  __device__ __forceinline__ void atomic_add(int i, double ax, double ay, double az) const {
    double* r = p + 4 * static_cast<int64_t>(i);
    unsafeAtomicAdd(r, ax);
    unsafeAtomicAdd(r + 1, ay);
    unsafeAtomicAdd(r + 2, az);
  }
};

struct MomAos3 {
  double* __restrict__ p;
  __device__ __forceinline__ void add(int i, double ax, double ay, double az) const {
    double* r = p + 3 * static_cast<int64_t>(i);
    r[0] += ax;
    r[1] += ay;
    r[2] += az;
  }
  __device__ __forceinline__ void atomic_add(int i, double ax, double ay, double az) const {
    double* r = p + 3 * static_cast<int64_t>(i);
    unsafeAtomicAdd(r, ax);
    unsafeAtomicAdd(r + 1, ay);
    unsafeAtomicAdd(r + 2, az);
  }
};

struct MomSoa {
  double* __restrict__ x;
  double* __restrict__ y;
  double* __restrict__ z;
  __device__ __forceinline__ void add(int i, double ax, double ay, double az) const {
    x[i] += ax;
    y[i] += ay;
    z[i] += az;
  }
  __device__ __forceinline__ void atomic_add(int i, double ax, double ay, double az) const {
    unsafeAtomicAdd(x + i, ax);
    unsafeAtomicAdd(y + i, ay);
    unsafeAtomicAdd(z + i, az);
  }
};

// Pair constants: c24 = 24*dt and c48 = 48*dt fold the time step into the
// force scalar; cl2 is the squared cutoff (pairs with r2 > cl2 contribute
// nothing, r2 == cl2 is kept, as in the reference's `if (r2 > CL2)`), rc its
// square root (computed on the host, so no kernel spends VALU on it).
struct PairParams {
  double c24, c48, cl2, box, inv_box, rc;
  float c24f, c48f;
};

// Minimum image in a periodic box: d - box * rint(d / box). For positions
// inside [0, box) the shift is -1, 0 or +1 box lengths and the result is the
// same double as the oracle's threshold form (d > box/2 ? d - box : ...).
template <bool PBC>
__device__ __forceinline__ void min_image(double& dx, double& dy, double& dz,
                                          const PairParams& pp) {
  if (PBC) {
    dx = fma(-pp.box, rint(dx * pp.inv_box), dx);
    dy = fma(-pp.box, rint(dy * pp.inv_box), dy);
    dz = fma(-pp.box, rint(dz * pp.inv_box), dz);
  }
}

// 1/x in FP64: the hardware reciprocal estimate refined by one
// Newton-Raphson step r <- r + r*(1 - x*r) (quadratic convergence). Measured
// on gfx950 over x in [0.4, 10]: the estimate alone is good to about 2^-24
// relative, one step to 11 ulp (2.4e-15), two steps to 0.5 ulp. One step
// keeps pair forces within ~1e-14 of the divide, far inside the 1e-12
// parity bar. Valid for the finite, normal x of a pair distance.
__device__ __forceinline__ double rcp_f64(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(fma(-x, r, 1.0), r, r);
}

// The reciprocals of x[0..N) in place (N = 1, 2 or 4) from one refined
// reciprocal of their product (Montgomery's simultaneous inversion):
// 1/x0 = x1 / (x0 x1) and so on. A refined reciprocal costs the
// transcendental unit's quarter rate plus two FMAs; four of them cost
// 96 cycles per wavefront, one of the product with the nine products
// around it 60. Each result carries five roundings more than the refined
// reciprocal's own error (about 11 ulp): about 16 ulp, 4e-15 relative. The
// factors are squared distances of LJ pairs, so their product stays far
// from overflow and underflow.
template <int N>
__device__ __forceinline__ void rcp_f64_batch(double (&x)[N]) {
  static_assert(N == 1 || N == 2 || N == 4, "batched reciprocal of 1, 2 or 4 values");
  if constexpr (N == 1) {
    x[0] = rcp_f64(x[0]);
  } else if constexpr (N == 2) {
    const double r = rcp_f64(x[0] * x[1]);
    const double a = x[0];
    x[0] = x[1] * r;
    x[1] = a * r;
  } else {
    const double p01 = x[0] * x[1], p23 = x[2] * x[3];
    const double r = rcp_f64(p01 * p23);
    const double r01 = r * p23, r23 = r * p01;
    const double a = x[0], c = x[2];
    x[0] = x[1] * r01;
    x[1] = a * r01;
    x[2] = x[3] * r23;
    x[3] = c * r23;
  }
}

// FP64 force scalar times dt: ((24 r^6 - 48) / r^14) * dt rewritten with
// s = 1/r2 as s^4 * (24 - 48 s^3) * dt (one reciprocal, no divide).
// Beyond the cutoff only the high word of s is cleared: s becomes a
// denormal of at most 2^-1042 whose third and fourth powers round to exactly
// 0, so df is exactly 0 at the cost of one 32-bit select.
__device__ __forceinline__ double lj_df_f64_s(double s, double r2, const PairParams& pp) {
  const int hi = (r2 > pp.cl2) ? 0 : __double2hiint(s);
  s = __hiloint2double(hi, __double2loint(s));
  const double s3 = s * s * s;
  return (s3 * s) * fma(-pp.c48, s3, pp.c24);
}

__device__ __forceinline__ double lj_df_f64(double r2, const PairParams& pp) {
  return lj_df_f64_s(rcp_f64(r2), r2, pp);
}

// FP64 force scalar without the factor 24*dt: s^4 (1 - 2 s^3) with s = 1/r2,
// exactly 0 where `inside` is false (only the high word of s is cleared, as
// in lj_df_f64). The tiled kernels multiply each row's sum by 24*dt once, so
// the constants here are instruction operands instead of registers.
__device__ __forceinline__ double lj_df_unit_f64(double s, bool inside) {
  const int hi = inside ? __double2hiint(s) : 0;
  s = __hiloint2double(hi, __double2loint(s));
  const double s3 = s * s * s;
  return (s3 * s) * fma(-2.0, s3, 1.0);
}

// Mixed-precision force scalar: r2 (computed in FP64) rounded once to FP32,
// then the reference's form (24 r^6 - 48) / (r^6 r^6 r^2) in FP32 with the
// hardware FP32 reciprocal (within 1 ulp) and no refinement: a reciprocal
// off by one ulp in the same direction for every pair moves the force by at
// most 5.2e-7 of max|F| at density 1 and 8.5e-7 at 0.5 (numpy study at
// N = 119,164), inside the 1e-6 / 2e-6 bounds. The cutoff test uses the
// FP64 r2 so the pair set is the same as in the FP64 kernels.
__device__ __forceinline__ float lj_df_mixed(double r2, const PairParams& pp) {
  const float r2f = static_cast<float>(r2);
  const float r6 = r2f * r2f * r2f;
  const float r14 = r6 * r6 * r2f;
  const float inv = __builtin_amdgcn_rcpf(r14);
  const float df = fmaf(pp.c24f, r6, -pp.c48f) * inv;
  return (r2 > pp.cl2) ? 0.0f : df;
}

// Kahan-compensated FP32 running sum.
struct KahanF32 {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const {
    return static_cast<double>(s) - static_cast<double>(c);
  }
};

// Sum over the `W` consecutive lanes of a wavefront segment (W a power of two
// <= 64); every lane of the segment receives the total.
template <int W>
__device__ __forceinline__ double segment_sum(double v) {
#pragma unroll
  for (int m = W / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, W);
  return v;
}

}  // namespace ljgpu
