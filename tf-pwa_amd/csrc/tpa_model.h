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
  int32_t shape;                     // bits 0-1: 0 none (1), 1 BW, 2 BWR; bit 2: times g0
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

// ---------------------------------------------------------------- four-body final states
// A decay group of four spin-0 finals is lowered to trees of three two-body vertices:
// vertex 0 is the top decay, vertices 1 and 2 the decays of the two resonances. Quasi-two-body
// trees (A -> R1 R2, R1 -> f f, R2 -> f f) have both resonance vertices under vertex 0;
// sequential trees (A -> R1 f, R1 -> R2 f, R2 -> f f) have vertex 2 under vertex 1.
// Partial waves: k = k0(chain) + (ls0 * n_ls1 + ls1) * n_ls2 + ls2.
#define TPA4_MAX_FINAL 4
#define TPA4_MAX_TOPO 8
#define TPA4_MAX_CHAIN 32

// Angle cache fields per four-body topology: for vertex v of the representative tree,
// 5v + 0: invariant mass of the core, 5v + 1..2: (alpha, beta) of outs[0] and 5v + 3..4:
// (alpha, beta) of outs[1] in the core's helicity frame. After all topologies: m_top, then the
// data mass of every final.
#define TPA4_NF_TOPO 15

typedef struct Tpa4Topo {
  int32_t core[3];        // finals below each vertex's core (bit mask; vertex 0: all)
  int32_t out[3][2];      // finals below each out of the representative decay (bit masks)
  int32_t parent[3];      // vertex whose out is this vertex's core (-1 for vertex 0)
  int32_t parent_out[3];  // which out (0 or 1) of the parent vertex
} Tpa4Topo;

// Node slots of a chain: 0 the top, 1 and 2 the resonances decaying at vertices 1 and 2,
// 3 + f final f (spin 0, helicity 0).
typedef struct Tpa4Chain {
  int32_t topo;
  int32_t swap[3];             // the chain lists vertex v's outs in the opposite order
  int32_t out_slot[3][2];      // node slot of each out, in this chain's order
  int32_t n_ls[3];
  int32_t l[3][TPA_MAX_LS];
  int32_t act[3][TPA_MAX_LS];  // 1: barrier uses 1/sqrt(P_l(z)); 0: clamped
  int32_t cg_off[3];           // CG tables [ls][i_out0][i_out1] in this chain's order
  int32_t k0;
  int32_t r_j2[2], r_n[2];     // resonances of slots 1 and 2
  int32_t r_hel[2][TPA_MAX_HEL];
  int32_t r_shape[2], r_bw_l[2];   // lineshape codes as TpaChain.shape
  double r_m0[2], r_g0[2];
  double r_q0m[2][2];          // nominal daughter masses of each resonance (q0 of its width)
  double d;                    // barrier radius parameter
} Tpa4Chain;

typedef struct TpaModel4 {
  int32_t n_final, n_topo, n_chain, P, lam, lam_pad;
  int32_t top_j2, n_top;
  int32_t hel_top[TPA_MAX_HEL];
  int32_t random_z;
  int32_t n_cg;
  Tpa4Topo topo[TPA4_MAX_TOPO];
  Tpa4Chain chain[TPA4_MAX_CHAIN];
} TpaModel4;
