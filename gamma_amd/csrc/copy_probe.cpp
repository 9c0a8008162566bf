// copy_probe.cpp -- calibration kernel for HBM traffic counters.
//
// Copies n doubles (default 2^27 = 1 GiB read + 1 GiB write) once per launch,
// so that FETCH_SIZE / WRITE_SIZE read under rocprofv3 --pmc for this kernel
// give the counter's scale against a known byte count, and its HIP-event time
// gives a measured streaming bandwidth next to the 8 TB/s specification.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

__global__ void __launch_bounds__(256) copy_f64(const double* __restrict__ a, double* __restrict__ b,
                                                long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
  long long n = argc > 1 ? atoll(argv[1]) : (1ll << 27);
  int reps = argc > 2 ? atoi(argv[2]) : 5;
  double *a, *b;
  if (hipMalloc(&a, n * sizeof(double)) != hipSuccess || hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
    fprintf(stderr, "allocation failed\n");
    return 1;
  }
  hipMemset(a, 0, n * sizeof(double));
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int blocks = 256 * 8 * 4;
  copy_f64<<<blocks, 256>>>(a, b, n);
  hipDeviceSynchronize();
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    hipEventRecord(e0);
    copy_f64<<<blocks, 256>>>(a, b, n);
    hipEventRecord(e1);
    hipEventSynchronize(e1);
    float ms;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  double bytes = 2.0 * n * sizeof(double);
  printf("{\"copy_bytes\": %.0f, \"best_ms\": %.4f, \"GBs\": %.1f}\n", bytes, best, bytes / (best * 1e-3) / 1e9);
  hipFree(a);
  hipFree(b);
  return 0;
}
