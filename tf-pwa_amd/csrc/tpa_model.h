// Model description shared by the host (filled through ctypes) and the kernels.
//
// A config.yml decay group of a three-body final state is lowered to this flat form:
// topologies (which pair of finals forms the resonance), chains (one resonance each, with
// its LS tables, CG offsets and lineshape) and the partial-wave index space
// k = k0(chain) + ls1 * n_ls2 + ls2. Every spin and helicity is stored doubled (2J, 2*lambda)
// so it is an exact integer.
#pragma once
#include <stdint.h>

#define TPA_MAX_FINAL 3
#define TPA_MAX_TOPO 3
#define TPA_MAX_CHAIN 32
#define TPA_MAX_LS 8
#define TPA_MAX_HEL 8
#define TPA_MAX_J2 7
#define TPA_MAX_P 128
#define TPA_MAX_LAM 64
#define TPA_MAX_CG 8192

// Angle cache fields per topology (the layout of the per-event angle SoA):
//   0: m_R, 1-2: (alpha, beta) of R in the top rest frame, 3-4: (alpha, beta) of the
//   bachelor X, 5-6 and 7-8: (alpha, beta) of the resonance daughters y and z in the R rest
//   frame, 9 + 3f .. 11 + 3f: aligned Euler angles (alpha, beta, gamma) of final f.
// After all topologies: m_top, then the data mass of every final.
#define TPA_NF_TOPO 18
#define TPA_F_MR 0
#define TPA_F_AR 1
#define TPA_F_AX 3
#define TPA_F_AY 5
#define TPA_F_AZ 7
#define TPA_F_AL 9

typedef struct TpaTopo {
  int32_t x, y, z;                   // finals: bachelor X and the resonance daughters (y, z)
                                     // in the representative chain's order
  int32_t r_first;                   // representative top decay lists the resonance first
  int32_t aligned[TPA_MAX_FINAL];    // final f carries an aligned angle in this topology
} TpaTopo;

typedef struct TpaChain {
  int32_t topo;
  int32_t top_r_first;               // this chain's top decay lists the resonance first
  int32_t r_y_first;                 // this chain's resonance decay lists topo.y first
  int32_t jr2;                       // 2J of the resonance
  int32_t n_r;                       // number of resonance helicities (2J+1)
  int32_t hel_r[TPA_MAX_HEL];        // 2*lambda of the resonance
  int32_t n_ls1, n_ls2;              // LS couplings of the production / decay vertex
  int32_t l1[TPA_MAX_LS], l2[TPA_MAX_LS];
  int32_t act1[TPA_MAX_LS], act2[TPA_MAX_LS];  // 1: barrier uses 1/sqrt(P_l(z)); 0: clamped
  int32_t k0;                        // first partial-wave index of the chain
  int32_t shape;                     // 0: none (1), 1: BW, 2: BWR
  int32_t bw_l;                      // orbital L of the running width
  int32_t cg1_off, cg2_off;          // CG tables [ls][i_b][i_c] in this chain's out order
  double m0, g0;                     // lineshape mass and width (folded builds)
  double m_y, m_z;                   // nominal daughter masses (q0 of the running width)
  double d;                          // barrier radius parameter
} TpaChain;

typedef struct TpaModel {
  int32_t n_final, n_topo, n_chain, P, lam, lam_pad;
  int32_t top_j2, n_top;
  int32_t hel_top[TPA_MAX_HEL];
  int32_t final_j2[TPA_MAX_FINAL];
  int32_t n_fh[TPA_MAX_FINAL];
  int32_t hel_f[TPA_MAX_FINAL][TPA_MAX_HEL];
  int32_t ref_topo[TPA_MAX_FINAL];   // topology whose frame defines each final's helicity
  int32_t random_z;
  int32_t n_cg;
  TpaTopo topo[TPA_MAX_TOPO];
  TpaChain chain[TPA_MAX_CHAIN];
} TpaModel;
