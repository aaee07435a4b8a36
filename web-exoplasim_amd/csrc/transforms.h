// Job lists of the spectral transform kernels. A job names arena slots by
// ExoSlot index, so one job list serves every member of a batch.
//
// A transform runs in two launches, as the reference's sp2fc/fc2gp and
// gp2fc/fc2sp pairs do, with the Fourier coefficients in the arena's
// Fourier slots between them:
//   inverse: k_leg_inv (one workgroup per zonal wavenumber m and job,
//            kernels_legendre.hip) then k_dft_inv (real FFTs, one latitude
//            row per lane, kernels_dft.hip);
//   forward: k_dft_fwd (real FFTs) then k_leg_fwd (per m).
#pragma once
#include "device.h"

enum InvKind {
  INV_PLAIN = 0,    // sp2fc with `table`: src0 -> four0; fc2gp: four0 -> dst0
  INV_DLAMBDA = 1,  // as PLAIN, and d/dlambda of four0 synthesised into dst1
  INV_DV2UV = 2,    // dv2uv: (D src0, zeta src1) -> (U four0, V four1) -> (dst0, dst1)
};

struct InvJob {
  int kind;
  int table;        // ExoTable of the Legendre table (PLAIN, DLAMBDA)
  int src0, src1;   // spectral slots
  int four0, four1; // Fourier slots
  int dst0, dst1;   // grid slots
  int per_level;    // 1: NLEV planes, 0: one plane
  int plavor;       // DV2UV: 1 removes the planetary vorticity from mode 1
};

#define EXO_MAX_JOBS 12

struct InvJobs {
  int n;
  InvJob j[EXO_MAX_JOBS];
};

// gp2fc: grid slot -> Fourier slot, per row and plane.
struct FwdJob {
  int src, dst, per_level;
};
struct FwdJobs {
  int n;
  FwdJob j[EXO_MAX_JOBS];
};

enum LegKind {
  LEG_PLAIN = 0,  // fc2sp with qc: src -> dst0
  LEG_UV2DV = 1,  // uv2dv (+kinetic energy when ke >= 0): (u, v) -> (D dst0, zeta dst1)
  LEG_QTEND = 2,  // flux-form tendency: (u, v, src) -> dst0
};
struct LegJob {
  int kind;
  int u, v, ke, src;
  int dst0, dst1;
  int per_level;
};
struct LegJobs {
  int n;
  LegJob j[EXO_MAX_JOBS];
};

static inline int job_planes(int per_level, int nlev) { return per_level ? nlev : 1; }
