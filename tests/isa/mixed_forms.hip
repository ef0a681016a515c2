// Compile-only probe (never launched): the inner loop of four pair forms,
// each over a row of 16-bit LDS offsets, four entries per iteration as in
// the tile kernel's batches, so that tests/test_mixed_forms_isa.py can count
// the VALU instructions per list entry each form needs on gfx950.
//   k_f64      the FP64 tile kernel's form (unit force scalar, 24-B entries)
//   k_current  the mixed kernel's form (FP64 differences and r2, the FP32
//              scalar (24 r^6 - 48) / r^14 with the hardware reciprocal)
//   k_int32_exact_cut  int32 fixed-point tile-relative coordinates (16-B
//              entries; exact integer differences), exact int64 r2 for the
//              cutoff, FP32 r2 and scalar, FP64 accumulation
//   k_int32_f32_cut    the same with the cutoff decided on the FP32 r2
// Fixed-point scaling (2^-27 per coordinate) is folded into one multiply of
// r2 and into the row total, as a kernel would.
#include <hip/hip_runtime.h>
#include <cstdint>

extern __shared__ char lds[];

__device__ __forceinline__ float df32(float r2) {
  const float r6 = r2 * r2 * r2;
  const float r14 = r6 * r6 * r2;
  const float inv = __builtin_amdgcn_rcpf(r14);
  return fmaf(24.0f, r6, -48.0f) * inv;
}

// Current form: FP64 positions (24 B), FP64 differences and r2.
__global__ void k_current(const uint16_t* __restrict__ offs, int np, double* out, double cl2) {
  const double xi = out[0], yi = out[1], zi = out[2];
  double ax = 0, ay = 0, az = 0;
  for (int k = threadIdx.x * 4; k < np; k += 256) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t a = offs[k + u];
      const double x = *reinterpret_cast<const double*>(lds + a);
      const double y = *reinterpret_cast<const double*>(lds + a + 8);
      const double z = *reinterpret_cast<const double*>(lds + a + 16);
      const double dx = x - xi, dy = y - yi, dz = z - zi;
      const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
      float df = df32(static_cast<float>(r2));
      df = r2 > cl2 ? 0.0f : df;
      const double d = df;
      ax = fma(d, dx, ax); ay = fma(d, dy, ay); az = fma(d, dz, az);
    }
  }
  out[threadIdx.x] = ax + ay + az;
}

// int32 fixed point (16-B int4 entries), exact int64 r2 for the cutoff,
// FP32 r2 for the force, FP64 accumulation of df * d.
__global__ void k_int32_exact_cut(const uint16_t* __restrict__ offs, int np, double* out,
                                  long long cl2q) {
  const int xi = reinterpret_cast<int*>(out)[0], yi = reinterpret_cast<int*>(out)[1],
            zi = reinterpret_cast<int*>(out)[2];
  double ax = 0, ay = 0, az = 0;
  for (int k = threadIdx.x * 4; k < np; k += 256) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t a = offs[k + u];
      const int4 p = *reinterpret_cast<const int4*>(lds + a);
      const int dx = p.x - xi, dy = p.y - yi, dz = p.z - zi;
      const long long r2q = (long long)dx * dx + (long long)dy * dy + (long long)dz * dz;
      const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
      const float r2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx)) * 0x1p-54f;
      float df = df32(r2);
      df = r2q > cl2q ? 0.0f : df;
      const double d = df;
      ax = fma(d, (double)fx, ax); ay = fma(d, (double)fy, ay); az = fma(d, (double)fz, az);
    }
  }
  out[threadIdx.x] = (ax + ay + az) * 0x1p-27;
}

// int32 fixed point, cutoff on the FP32 r2 (flips pairs at r = 3).
__global__ void k_int32_f32_cut(const uint16_t* __restrict__ offs, int np, double* out, float cl2) {
  const int xi = reinterpret_cast<int*>(out)[0], yi = reinterpret_cast<int*>(out)[1],
            zi = reinterpret_cast<int*>(out)[2];
  double ax = 0, ay = 0, az = 0;
  for (int k = threadIdx.x * 4; k < np; k += 256) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t a = offs[k + u];
      const int4 p = *reinterpret_cast<const int4*>(lds + a);
      const int dx = p.x - xi, dy = p.y - yi, dz = p.z - zi;
      const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
      const float r2 = fmaf(fz, fz, fmaf(fy, fy, fx * fx)) * 0x1p-54f;
      float df = df32(r2);
      df = r2 > cl2 ? 0.0f : df;
      const double d = df;
      ax = fma(d, (double)fx, ax); ay = fma(d, (double)fy, ay); az = fma(d, (double)fz, az);
    }
  }
  out[threadIdx.x] = (ax + ay + az) * 0x1p-27;
}

// FP64 form for comparison (unit scalar, as the FP64 tile kernel).
__global__ void k_f64(const uint16_t* __restrict__ offs, int np, double* out, double cl2) {
  const double xi = out[0], yi = out[1], zi = out[2];
  double ax = 0, ay = 0, az = 0;
  for (int k = threadIdx.x * 4; k < np; k += 256) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const uint32_t a = offs[k + u];
      const double x = *reinterpret_cast<const double*>(lds + a);
      const double y = *reinterpret_cast<const double*>(lds + a + 8);
      const double z = *reinterpret_cast<const double*>(lds + a + 16);
      const double dx = x - xi, dy = y - yi, dz = z - zi;
      const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
      const double r = __builtin_amdgcn_rcp(r2);
      double s = fma(fma(-r2, r, 1.0), r, r);
      const int hi = (r2 > cl2) ? 0 : __double2hiint(s);
      s = __hiloint2double(hi, __double2loint(s));
      const double s3 = s * s * s;
      const double d = (s3 * s) * fma(-2.0, s3, 1.0);
      ax = fma(d, dx, ax); ay = fma(d, dy, ay); az = fma(d, dz, az);
    }
  }
  out[threadIdx.x] = ax + ay + az;
}
