// Instances and launcher of the z-march apply.
#include "zmarch.h"

namespace octhip {

namespace {
template <int LPT, int LC, bool NL>
int launch_nl(const ZArgs& z, int epi, int grid, size_t shmem, hipStream_t s) {
  constexpr int NT = ZNC * LC / LPT;
  if constexpr (NT > 1024) {
    return fail(OCTHIP_ERR_INVALID, "z-march: %d lanes per workgroup exceed 1024 threads", LC);
  } else if (epi == EPI_TAYLOR) {
    if constexpr (LC >= 2) {
      hipLaunchKernelGGL((k_zmarch<LPT, LC, EPI_TAYLOR, NL>), dim3(grid), dim3(NT), shmem, s, z);
    } else {
      return fail(OCTHIP_ERR_INVALID, "z-march Taylor epilogue needs complex lane pairs");
    }
  } else if (epi == EPI_TAYLOR_X) {
    if constexpr (LC >= 2) {
      hipLaunchKernelGGL((k_zmarch<LPT, LC, EPI_TAYLOR_X, NL>), dim3(grid), dim3(NT), shmem, s, z);
    } else {
      return fail(OCTHIP_ERR_INVALID, "z-march Taylor epilogue needs complex lane pairs");
    }
  } else {
    hipLaunchKernelGGL((k_zmarch<LPT, LC, EPI_STORE, NL>), dim3(grid), dim3(NT), shmem, s, z);
  }
  return OCTHIP_SUCCESS;
}

template <int LPT, int LC>
int launch_lc(const ZArgs& z, int epi, int grid, int zseg, hipStream_t s) {
  const size_t shmem = zmarch_shmem(LC, zseg);
  ProfScope prof("k_zmarch", s);
  const int rc = z.a.kslot ? launch_nl<LPT, LC, true>(z, epi, grid, shmem, s)
                         : launch_nl<LPT, LC, false>(z, epi, grid, shmem, s);
  if (rc) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_zmarch launch: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}
template <int LPT, int LC>
int launch_multi_lc(const ZArgs* zs, const int4* sched, int len, int epi, bool nl, int zseg, hipStream_t s) {
  constexpr int NT = ZNC * LC / LPT;
  const size_t shmem = zmarch_shmem(LC, zseg);
  ProfScope prof("k_zmarch", s);
  if constexpr (NT > 1024) {
    return fail(OCTHIP_ERR_INVALID, "z-march: %d lanes per workgroup exceed 1024 threads", LC);
  } else if (epi == EPI_TAYLOR_X) {
    return fail(OCTHIP_ERR_UNSUPPORTED, "batched z-march: no psib2/density epilogue");
  } else if (epi == EPI_TAYLOR) {
    if constexpr (LC >= 2) {
      if (nl) hipLaunchKernelGGL((k_zmarch_multi<LPT, LC, EPI_TAYLOR, true>), dim3(len), dim3(NT), shmem, s, zs, sched);
      else hipLaunchKernelGGL((k_zmarch_multi<LPT, LC, EPI_TAYLOR, false>), dim3(len), dim3(NT), shmem, s, zs, sched);
    } else {
      return fail(OCTHIP_ERR_INVALID, "z-march Taylor epilogue needs complex lane pairs");
    }
  } else {
    if (nl) hipLaunchKernelGGL((k_zmarch_multi<LPT, LC, EPI_STORE, true>), dim3(len), dim3(NT), shmem, s, zs, sched);
    else hipLaunchKernelGGL((k_zmarch_multi<LPT, LC, EPI_STORE, false>), dim3(len), dim3(NT), shmem, s, zs, sched);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OCTHIP_ERR_HIP, "k_zmarch_multi launch: %s", hipGetErrorString(e));
  return OCTHIP_SUCCESS;
}
template <int LPT, int LC>
const void* kernel_of(int epi, bool nl, bool multi) {
  if constexpr (ZNC * LC / LPT > 1024) {
    return nullptr;
  } else if (epi == EPI_TAYLOR_X) {
    if constexpr (LC >= 2) {
      if (multi) return nullptr;
      return nl ? reinterpret_cast<const void*>(&k_zmarch<LPT, LC, EPI_TAYLOR_X, true>)
                : reinterpret_cast<const void*>(&k_zmarch<LPT, LC, EPI_TAYLOR_X, false>);
    } else {
      return nullptr;
    }
  } else if (epi == EPI_TAYLOR) {
    if constexpr (LC >= 2) {
      if (multi)
        return nl ? reinterpret_cast<const void*>(&k_zmarch_multi<LPT, LC, EPI_TAYLOR, true>)
                  : reinterpret_cast<const void*>(&k_zmarch_multi<LPT, LC, EPI_TAYLOR, false>);
      return nl ? reinterpret_cast<const void*>(&k_zmarch<LPT, LC, EPI_TAYLOR, true>)
                : reinterpret_cast<const void*>(&k_zmarch<LPT, LC, EPI_TAYLOR, false>);
    } else {
      return nullptr;
    }
  }
  if (multi)
    return nl ? reinterpret_cast<const void*>(&k_zmarch_multi<LPT, LC, EPI_STORE, true>)
              : reinterpret_cast<const void*>(&k_zmarch_multi<LPT, LC, EPI_STORE, false>);
  return nl ? reinterpret_cast<const void*>(&k_zmarch<LPT, LC, EPI_STORE, true>)
            : reinterpret_cast<const void*>(&k_zmarch<LPT, LC, EPI_STORE, false>);
}
}  // namespace

// Workgroups of one z-march instance (k_zmarch, or k_zmarch_multi when multi) that its
// registers alone let run on a CU (LDS not counted): the segment-length choice keeps the
// LDS of that many resident.  Cached per instance; 1 for shapes without an instance.
int zmarch_reg_blocks(int lc, int lpt, int epi, bool nl, bool multi) {
  static int cache[2][2][6][3][2] = {};  // [multi][lpt-1][log2 lc][epi][nl], 0 = not asked yet
  int lg = 0;
  while ((1 << lg) < lc) ++lg;
  if (lpt < 1 || lpt > 2 || lg > 5 || (1 << lg) != lc || epi < 0 || epi > 2) return 1;
  int& c = cache[multi ? 1 : 0][lpt - 1][lg][epi][nl ? 1 : 0];
  if (c > 0) return c;
  const void* f = nullptr;
  if (lpt == 2) {
    switch (lc) {
      case 2: f = kernel_of<2, 2>(epi, nl, multi); break;
      case 4: f = kernel_of<2, 4>(epi, nl, multi); break;
      case 8: f = kernel_of<2, 8>(epi, nl, multi); break;
      case 16: f = kernel_of<2, 16>(epi, nl, multi); break;
      case 32: f = multi ? nullptr : kernel_of<2, 32>(epi, nl, false); break;
      default: break;
    }
  } else if (!multi) {
    switch (lc) {
      case 1: f = kernel_of<1, 1>(epi, nl, false); break;
      case 2: f = kernel_of<1, 2>(epi, nl, false); break;
      case 4: f = kernel_of<1, 4>(epi, nl, false); break;
      case 8: f = kernel_of<1, 8>(epi, nl, false); break;
      case 16: f = kernel_of<1, 16>(epi, nl, false); break;
      default: break;
    }
  }
  int nb = 1;
  if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, ZNC * lc / lpt, 0) != hipSuccess || nb < 1) nb = 1;
  c = nb;
  return nb;
}

int launch_zmarch_multi(const ZArgs* zs, const int4* sched, int len, int lc, int lpt, int epi, bool nl, int zseg,
                        hipStream_t s) {
  if (len <= 0) return OCTHIP_SUCCESS;
  if (zmarch_shmem(lc, zseg) > 160 * 1024)
    return fail(OCTHIP_ERR_INVALID, "z-march: lane chunk %d with z segment %d needs more than 160 KB of LDS", lc, zseg);
  if (lpt == 2) {
    switch (lc) {
      case 2: return launch_multi_lc<2, 2>(zs, sched, len, epi, nl, zseg, s);
      case 4: return launch_multi_lc<2, 4>(zs, sched, len, epi, nl, zseg, s);
      case 8: return launch_multi_lc<2, 8>(zs, sched, len, epi, nl, zseg, s);
      case 16: return launch_multi_lc<2, 16>(zs, sched, len, epi, nl, zseg, s);
      default: break;
    }
  }
  return fail(OCTHIP_ERR_INVALID, "no multi z-march instance for lane chunk %d, %d lanes per thread", lc, lpt);
}

int launch_zmarch(const ZArgs& z, int lc, int lpt, int epi, int grid, int zseg, hipStream_t s) {
  if (lpt == 2) {
    switch (lc) {
      case 2: return launch_lc<2, 2>(z, epi, grid, zseg, s);
      case 4: return launch_lc<2, 4>(z, epi, grid, zseg, s);
      case 8: return launch_lc<2, 8>(z, epi, grid, zseg, s);
      case 16: return launch_lc<2, 16>(z, epi, grid, zseg, s);
      case 32: return launch_lc<2, 32>(z, epi, grid, zseg, s);
      default: break;
    }
  } else {
    switch (lc) {
      case 1: return launch_lc<1, 1>(z, epi, grid, zseg, s);
      case 2: return launch_lc<1, 2>(z, epi, grid, zseg, s);
      case 4: return launch_lc<1, 4>(z, epi, grid, zseg, s);
      case 8: return launch_lc<1, 8>(z, epi, grid, zseg, s);
      case 16: return launch_lc<1, 16>(z, epi, grid, zseg, s);
      default: break;
    }
  }
  return fail(OCTHIP_ERR_INVALID, "no z-march instance for lane chunk %d, %d lanes per thread", lc, lpt);
}

}  // namespace octhip
