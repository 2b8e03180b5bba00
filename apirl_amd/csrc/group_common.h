// Device helpers of the grouped (multi-job) launches: the next virtual blocks a
// workgroup takes, and the job that owns a virtual block.
// Included inside namespace apirl::(anonymous) of the kernel sources.
#ifndef APIRL_GROUP_COMMON_H
#define APIRL_GROUP_COMMON_H

// Virtual blocks are handed out by a per-phase counter, `grab` at a time (the round
// header's, tuned per batch; 1 if unset), so a workgroup that drew cheap tiles takes
// more (gather tiles differ ~10x in cost).  Returns the first virtual block of the
// grab (>= total: no work left).
__device__ __forceinline__ int grab_size(const BatchHeader* hdr) {
  return hdr->grab > 0 ? hdr->grab : 1;
}
__device__ __forceinline__ long grab_vb(unsigned long long* ctr, int grab) {
  __shared__ long s_vb;
  __syncthreads();  // every lane has read the previous grab
  if (threadIdx.x == 0) s_vb = (long)atomicAdd(ctr, (unsigned long long)grab);
  __syncthreads();
  return s_vb;
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
