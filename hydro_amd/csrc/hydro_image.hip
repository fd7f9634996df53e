// Density image of a grid, computed where the state lives: the reference's
// single-rank PPM image (HydroC99_2Domp/Src/Image.c:71-166, written by
// :291-347 through DumpImage :385-396) needs the maximum density over every
// padded cell and then one sample every `shrink` cells of the interior.  Here
// both run on the GPU and only the w*h*3 bytes of the image cross PCIe (the
// reference reads the whole state on the host).
#include "hydro_kernels.hpp"

namespace hydro {

// Order-preserving key of a double: unsigned order of keys is the order of
// the values (-0 below +0).  Key 0 belongs to a NaN bit pattern only, so it
// serves as "no value seen".
__device__ __forceinline__ uint64_t order_key(double x) {
  const uint64_t b = (uint64_t) __double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_value(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long) b);
}

// Maximum of the density plane over all (nx+4) x (ny+4) padded cells,
// NaN values skipped (Image.c:71-110 compares with `v > max`, which a NaN
// never passes).  A workgroup strides over rows; one atomic per wave.
template <typename S>
__global__ __launch_bounds__(256) void k_image_max(GridGeom g, const S* __restrict__ rho,
                                                   unsigned long long* key) {
  const int nxt = g.nx + 4, nyt = g.ny + 4;
  unsigned long long m = 0;
  for (int j = blockIdx.x; j < nyt; j += gridDim.x) {
    const S* row = rho + (long) j * g.pitch;
    for (int i = threadIdx.x; i < nxt; i += blockDim.x) {
      const double v = (double) row[i];
      if (v == v) {
        const unsigned long long k = order_key(v);
        m = k > m ? k : m;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m != 0) atomicMax(key, m);
}

// One thread per pixel: interior cell (2 + cx*shrink, 2 + cy*shrink) ->
// (uint8_t) fabs(255 * rho / max), the same byte in R, G and B (Image.c:150-
// 162).  The C cast is exact truncation for values below 256; for larger,
// infinite or NaN values (undefined in C) the byte follows the x86-64
// conversion the reference's gcc build performs (cvttsd2si to int32, low
// byte; 0x80000000 when out of the int32 range).  The maximum is NaN when
// the first padded cell, where the reference's scan starts, is NaN.
template <typename S>
__global__ __launch_bounds__(256) void k_image_pixels(GridGeom g, const S* __restrict__ rho,
                                                      const unsigned long long* key, int shrink,
                                                      int w, int h, unsigned char* rgb) {
  const long idx = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long) w * h) return;
  const int cy = (int) (idx / w), cx = (int) (idx - (long) cy * w);
  const double first = (double) rho[0];
  const unsigned long long k = *key;
  const double mx = (first != first || k == 0) ? __longlong_as_double(0x7ff8000000000000ll)
                                               : key_value(k);
  const double r = (double) rho[(long) (2 + cy * shrink) * g.pitch + 2 + cx * shrink];
  const double a = fabs(255.0 * r / mx);
  const unsigned char b = (a < 2147483648.0) ? (unsigned char) (int) a : (unsigned char) 0;
  rgb[3 * idx + 0] = b;
  rgb[3 * idx + 1] = b;
  rgb[3 * idx + 2] = b;
}

__global__ void k_image_clear(unsigned long long* key) { *key = 0; }

template <typename S>
void launch_image(const GridGeom& g, const S* rho, int shrink, int w, int h,
                  unsigned long long* key, unsigned char* rgb, hipStream_t st) {
  hipLaunchKernelGGL(k_image_clear, dim3(1), dim3(1), 0, st, key);
  const int blocks = g.ny + 4 < 2048 ? g.ny + 4 : 2048;
  hipLaunchKernelGGL(k_image_max<S>, dim3(blocks), dim3(256), 0, st, g, rho, key);
  const long n = (long) w * h;
  if (n > 0)
    hipLaunchKernelGGL(k_image_pixels<S>, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, st,
                       g, rho, key, shrink, w, h, rgb);
}
template void launch_image<double>(const GridGeom&, const double*, int, int, int,
                                   unsigned long long*, unsigned char*, hipStream_t);
template void launch_image<float>(const GridGeom&, const float*, int, int, int,
                                  unsigned long long*, unsigned char*, hipStream_t);

}  // namespace hydro
