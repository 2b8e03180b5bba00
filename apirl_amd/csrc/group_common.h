// Device helpers of the grouped (multi-job) launches: the share of a phase's
// virtual blocks that one workgroup takes, and the job that owns a virtual block.
// Included inside namespace apirl::(anonymous) of the kernel sources.
#ifndef APIRL_GROUP_COMMON_H
#define APIRL_GROUP_COMMON_H

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD b % 8); the
// remap gives each XCD one contiguous run of workgroups, and each workgroup one
// contiguous run of virtual blocks, so the blocks of one job (its image, tables and
// sinogram) stay in one or few XCDs' L2.
__device__ __forceinline__ void vb_range(long total, long& lo, long& hi) {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned xcd = b % 8, per = nb / 8, rem = nb % 8;
  const unsigned lb = xcd * per + (xcd < rem ? xcd : rem) + b / 8;
  const long chunk = (total + nb - 1) / nb;
  lo = (long)lb * chunk;
  hi = lo + chunk < total ? lo + chunk : total;
}

// Last job whose first virtual block is <= vb (jobs are sorted by first).
template <class J>
__device__ __forceinline__ int find_job(const J* jobs, int n, long vb) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].first <= vb) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#endif
