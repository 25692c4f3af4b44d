// gd_device.hpp -- data layout shared by the host orchestration and the HIP
// kernels of the barotropic step.
//
// HBM layout (per member, one arena): every 2-D field is `M` rows of `pitch`
// elements, longitude contiguous (pitch = N rounded up to 32 elements so each
// row starts on a 256-byte boundary).  Half-lat fields use rows 0..M-2.
// State slots:
//   X  - state at the start of a step: gd, U, V and u, v ("full" state)
//   Y,Z- outputs of integrator calls (ping-pong): gd, U, V; u and v are
//        derived on the fly exactly as update_state computes them
//        (dycore_mod.F90:636-645)
//   P1,P2 - the two predictor states of predict_correct
// Tendencies Tb (second predictor) and Tc (corrector) feed the beta inner
// products; du/dv are stored in the compute type, dgd always in FP64.
#pragma once

#include <cstdint>

namespace gd {

constexpr int kBlock = 256;   // threads of the row kernels
constexpr int kXcd = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin

enum Slot { SX = 0, SY = 1, SZ = 2, SP1 = 3, SP2 = 4, NSLOT = 5 };
enum Fld { FGD = 0, FU = 1, FV = 2, Fu = 3, Fv = 4, NFLD = 5 };
enum Comp { CDU = 0, CDV = 1, CDGD = 2 };
enum Phase { PH_A = 0, PH_B = 1, PH_C = 2, PH_T = 3, PH_D = 4, PH_P = 5 };

constexpr int kFiltRows = 8;     // staged rows per filter tile (+8 weight rows)
constexpr int kChunk = 64;       // slots (F1) / columns (F2) per filter workgroup

// Per full row j (0-based).  Half-row entries at index j describe half row j
// (0 for j = M-1).
struct RowCoef {
  double fcos, f, c, fdlon, fdlat;     // full row j (pole rows: reset cos)
  double hcos, hdlon, hdlat;           // half row j
  double k_uadv_lon, k_uadv_lat;       // 0.25/fdlon, 0.25/fdlat
  double k_vadv_lon, k_vadv_lat;       // 0.25/hdlon, 0.25/hdlat
  double c1, c2;                       // hcos[j-1]/fcos[j], hcos[j]/fcos[j]
  double fdlon2, fdlat2, hdlon2, hdlat2;
  double mass_w;                       // (fcos*dlon)*dlat
  int32_t stage[3];                    // staging slot of du/dv/dgd, -1 if none
  int32_t pad;
};

struct Member {
  double *st[NSLOT][NFLD];   // state fields (u/v read only from slot X on the step path)
  double *ghs;
  void *tb[3];               // Tb: du, dv (compute type), dgd (double)
  void *tc[3];               // Tc
  double *part;              // beta partials [M][3 comps][nchunk][2]
  double *dpart;             // diag partials [M][4]
  double *sx;                // staged raw filter rows [nslots][pitch]
  double *s1;                // [nslots] row inner product before filtering
  double *beta;              // [1] beta of the last integrator call
  double *div, *vor;         // history diagnostics
  double *diag_out;          // [cap][3]: mass, energy, beta per step
};

struct StageSlot {
  int32_t comp;   // Comp
  int32_t row;    // full row (du, dgd) or half row (dv)
  int32_t kept;   // kept FFTPACK slots (2c+2)
  int32_t pad;
};

// A filter tile: up to kFiltRows staged rows (member, slot) with one slot
// range.  Kept mode: projection onto slots [0, hi).  Complement mode:
// x minus the projection onto slots [lo, N).
struct FiltTile {
  int32_t nrows;
  int32_t complement;
  int32_t lo, hi;              // slot range handled by the GEMM
  int32_t fp64;                // FP64 GEMM (all tiles in FP64 mode; dgd tiles in mixed mode)
  int32_t pad;
  int32_t member[kFiltRows];
  int32_t slot[kFiltRows];
};

struct Ctrl {
  int32_t step;                // diag ring write position (all members)
  uint32_t done[1];            // per-member block counters follow (flexible)
};

struct Args {
  const Member *mem;
  const RowCoef *coef;
  const StageSlot *slots;
  const FiltTile *tiles;
  double *fcoef;               // filter coefficients [ntile][kFiltRows][kpad]
  double *fs2;                 // s2 partials [ntile][nchunk1][kFiltRows]
  const void *basis;           // B[s][l], slot-major, pitch lpad (compute type)
  const void *basis_t;         // Bt[l][s], longitude-major, pitch spad
  const void *basis64;         // FP64 copies (same as basis/basis_t in FP64 mode)
  const void *basis_t64;
  int lpad, spad;
  uint32_t *done;              // per-member finish counters
  int32_t *step_ctr;
  int N, M, pitch, nmem;
  int m0;                      // first member of a row launch
  int nseg;                    // longitude segments of kBlock points per row
  int ntiles;                  // work items of this launch
  int in_full;                 // k_diffuse: read stored u, v of s_in
  int nslots, nftiles, kpad, nchunk1, nchunk2, nchunk_row;
  int pass, phase;
  int s_in, s0, s_out;
  int upwind;
  int qcon;
  int diag_cap;
  double hdt;                  // dt/2 of a predictor phase
  double dt;                   // integrator time step (corrector)
  double kdiff;                // (sign*dt)*diffusion_coef
  double beta_lon, beta_lat;
  double dlon, dlat;
};

}  // namespace gd
