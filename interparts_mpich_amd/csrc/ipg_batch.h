// ipg_batch.h -- one launch of a kernel over the work of many contexts.
//
// Every kernel of the substep is written as a body that takes its arguments
// as one struct and its block index as a VB (the block's (x, y, z) and the
// grid it belongs to).  The single-context kernel passes the hardware block
// index; k_batch runs the same body over a merged index space in which case s
// owns the blocks [seg[s].b0, seg[s].b0 + gx*gy*gz), each case with its own
// argument struct.  So the cases of an ensemble advance through a substep
// stage in one launch per kernel (per FFT length, whose plans are compile-
// time), whatever their grids and sphere counts.
//
// Segments start at multiples of 8 blocks and a persistent grid is a multiple
// of 8 blocks, so a block's position modulo 8 -- the XCD the hardware hands it
// to -- is the same as in the case's own launch, and the XCD-aware tile orders
// of the kernels carry over.
#pragma once
#include <hip/hip_runtime.h>

struct VB {
  int x, y, z;     // block index
  int gx, gy, gz;  // grid of the launch it belongs to
};

struct BatchSeg {
  int b0;          // first merged block (a multiple of 8)
  int gx, gy, gz;  // the case's own grid
};

struct BatchMap {
  const BatchSeg *seg;        // [nseg]
  const int *of;              // [total / 8]: the segment of each run of 8 merged blocks
  int total;                  // merged blocks, a multiple of 8 (padding blocks do nothing)
};

// the merged kernels of one body: one workgroup per block, and persistent
struct BatchFns { const void *one, *persist; };

#if defined(__HIPCC__)
// (the kernels below are built by hipcc only; host-only builds of the headers,
// such as the index checks' g++ builds, see the structs above)
__device__ __forceinline__ VB vb_hw() {
  return VB{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}

// Block b of the merged space runs case of[b / 8]'s block b - b0, if that is
// inside the case's grid.
// The case's argument struct is read through the constant address space, as
// a kernel's own arguments are: its loads are then known to be invariant, so
// the compiler reloads a value where it is needed rather than holding it in
// registers (read as ordinary global memory, the Thomas solve's 220 VGPRs
// become 256 and its wavefronts per SIMD fall from 2 to 1).
#if defined(__HIP_DEVICE_COMPILE__)
#define IPG_CONST_AS __attribute__((address_space(4)))
#else
#define IPG_CONST_AS
#endif
template <class K>
__device__ __forceinline__ void batch_item(const typename K::Args *__restrict__ args, const BatchMap &m, int b) {
  typedef IPG_CONST_AS const typename K::Args CArgs;
  typedef IPG_CONST_AS const BatchSeg CSeg;
  typedef IPG_CONST_AS const int CInt;
  const int s = ((CInt *)m.of)[b >> 3];
  const BatchSeg sg = ((CSeg *)m.seg)[s];
  const int l = b - sg.b0;
  if (l < sg.gx * sg.gy * sg.gz) {
    const VB v{l % sg.gx, (l / sg.gx) % sg.gy, l / (sg.gx * sg.gy), sg.gx, sg.gy, sg.gz};
    K::run(*(const typename K::Args *)((CArgs *)args + s), v);
  }
}

// The body of a single-context kernel whose only argument is its argument
// struct: the struct is read where the kernel's arguments lie, through the
// constant address space, as batch_item reads a case's struct.  (Bound to a
// reference as an ordinary parameter, a large struct such as the IBM's is
// copied to scratch memory: 440-496 bytes per lane.)
template <class K>
__device__ __forceinline__ void run_single() {
  typedef IPG_CONST_AS const typename K::Args CArgs;
  K::run(*(const typename K::Args *)((CArgs *)__builtin_amdgcn_kernarg_segment_ptr()), vb_hw());
}

// Two forms of the merged launch, chosen per stage by the batch's tuner:
//  * one workgroup per merged block (the grid is the merged space), with the
//    registers of the single-context kernel;
//  * a persistent grid, smaller than the merged space, each workgroup
//    striding over it.  The barrier between items keeps one item's LDS reads
//    apart from the next item's writes; a body never leaves a whole wavefront
//    out of a barrier of its own (its out-of-range lanes stay in the loop or
//    skip their stores), so every wavefront meets the same barriers.  The loop
//    costs some registers (the RK stage's 119-125 VGPRs become 126-147).
// K: a kernel body (K::run), its argument struct K::Args, its workgroup size
// K::THREADS and K::WPE, the minimum waves per SIMD its single-context kernel
// asks for (0: none).
template <class K>
__global__ __launch_bounds__(K::THREADS) void k_batch(const typename K::Args *__restrict__ args, BatchMap m) {
  if ((int)blockIdx.x < m.total) batch_item<K>(args, m, blockIdx.x);
}
template <class K>
__global__ __launch_bounds__(K::THREADS) __attribute__((amdgpu_waves_per_eu(K::WPE)))
void k_batch_wpe(const typename K::Args *__restrict__ args, BatchMap m) {
  if ((int)blockIdx.x < m.total) batch_item<K>(args, m, blockIdx.x);
}
template <class K>
__global__ __launch_bounds__(K::THREADS) void k_batch_persist(const typename K::Args *__restrict__ args, BatchMap m) {
  for (int b = blockIdx.x; b < m.total; b += gridDim.x) {
    batch_item<K>(args, m, b);
    __syncthreads();
  }
}

template <class K>
inline BatchFns batch_kernels() {
  if constexpr (K::WPE > 0) return BatchFns{(const void *)k_batch_wpe<K>, (const void *)k_batch_persist<K>};
  else return BatchFns{(const void *)k_batch<K>, (const void *)k_batch_persist<K>};
}
#endif  // __HIPCC__
