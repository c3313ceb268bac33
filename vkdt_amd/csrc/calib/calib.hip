// calib.hip -- instruction-rate calibration kernels (measurement only, not
// part of the raw-development path; built as libvkdt_calib.so).
//
// Each kernel runs `iters` rounds of K independent dependency chains of one
// VALU instruction in every lane and stores the chains' sum, so nothing is
// dead.  With enough waves resident the chains hide the instruction latency
// and the kernel time gives the issue rate of that instruction:
//   wave-instructions = waves * iters * K (one per chain and round)
#include <hip/hip_runtime.h>
#include <stdint.h>

#define K 8

template<int OP>
__global__ void __launch_bounds__(256) k_calib(float *out, int iters, float a, float b)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  float c[K];
#pragma unroll
  for(int k = 0; k < K; k++) c[k] = 1e-3f * (float)(t + k);
  for(int it = 0; it < iters; it++)
  {
#pragma unroll
    for(int k = 0; k < K; k++)
    {
      if(OP == 0) c[k] = __builtin_fmaf(c[k], a, b);             // v_fma_f32
      else if(OP == 1) c[k] = c[k] * a;                           // v_mul_f32
      else if(OP == 2) c[k] = __builtin_amdgcn_exp2f(c[k]);       // v_exp_f32
      else if(OP == 3) c[k] = __builtin_amdgcn_rcpf(c[k]);        // v_rcp_f32
      else if(OP == 4) c[k] = __builtin_amdgcn_sqrtf(c[k]);       // v_sqrt_f32
      else if(OP == 5) c[k] = (k & 1) ? __builtin_amdgcn_exp2f(c[k]) : __builtin_fmaf(c[k], a, b);   // half each
      else if(OP == 6) c[k] = (k & 1) ? __int_as_float(__float_as_int(c[k]) + 1) : __builtin_fmaf(c[k], a, b);   // v_add_u32 / v_fma_f32
    }
  }
  float s = 0.0f;
#pragma unroll
  for(int k = 0; k < K; k++) s += c[k];
  out[t] = s;
}

// op: 0 fma, 1 mul, 2 exp, 3 rcp, 4 sqrt, 5 fma and exp chains alternating,
// 6 fma and integer add chains alternating.  out holds blocks * 256 floats.
// Returns 0, or -1 for an unknown op, or the HIP error code.
extern "C" int vkdt_calib_run(int op, float *out, int blocks, int iters, void *stream)
{
  const dim3 g(blocks), b(256);
  hipStream_t s = (hipStream_t)stream;
  switch(op)
  {
    case 0: k_calib<0><<<g, b, 0, s>>>(out, iters, 0.999f, 1e-4f); break;
    case 1: k_calib<1><<<g, b, 0, s>>>(out, iters, 0.9999f, 0.0f); break;
    case 2: k_calib<2><<<g, b, 0, s>>>(out, iters, 0.0f, 0.0f); break;
    case 3: k_calib<3><<<g, b, 0, s>>>(out, iters, 0.0f, 0.0f); break;
    case 4: k_calib<4><<<g, b, 0, s>>>(out, iters, 0.0f, 0.0f); break;
    case 5: k_calib<5><<<g, b, 0, s>>>(out, iters, 0.999f, 1e-4f); break;
    case 6: k_calib<6><<<g, b, 0, s>>>(out, iters, 0.999f, 1e-4f); break;
    default: return -1;
  }
  return (int)hipGetLastError();
}

// instructions per chain and round of each op (K chains per thread)
extern "C" int vkdt_calib_chains(void) { return K; }
