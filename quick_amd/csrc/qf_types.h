// Plain-old-data layouts shared by the host runtime and the kernels.
//
// A molecule's shell pairs are grouped into PairBlocks: one block per
// (pair class, primitive-pair count K).  Inside a block the pairs are sorted
// by Schwarz bound, descending, so the kets a bra can meet form a prefix.
// Primitive-pair data is stored field-major and lane-interleaved,
// prim[(field*K + k)*n + j], so 64 lanes that read 64 consecutive kets issue
// one coalesced load per field.
#pragma once

namespace qf {

constexpr int WAVE = 64;   // lanes per wavefront on CDNA

// Workgroups of a wave-per-quartet launch whose class keeps its node-program
// slots in global memory: the slot buffer holds this many slices.
constexpr int COOP_GLOBAL_SLICES = 1024;

// PF_ALPHA: exponent of the first Q-shell's primitive (the second's is
// zeta - alpha); read only by the gradient kernels.
enum PrimField { PF_ZETA = 0, PF_OOZ2, PF_PX, PF_PY, PF_PZ, PF_KC, PF_CUT, PF_ALPHA, PF_N };

struct PairInfo {
    int qa, qb;      // Q-shell ids (an SP shell is split into an s and a p Q-shell)
    int fa, fb;      // first basis function of the two Q-shells
    int sa, sb;      // original shell ids (density-screen matrix index)
    int qfs;         // Qfinal(orig a) + Qfinal(orig b): QUICK's NABCD contribution
    int id;          // unique pair id within the molecule
    double ycut;     // Schwarz bound of the original shell pair
    double A[3];     // centre of Q-shell a
    double AB[3];    // A - B
    int cola, colb;  // first column of Q-shells a and b in the block's compact column space
};

struct PairBlock {
    const PairInfo* info;   // n entries
    const double* prim;     // PF_N * K * n
    const int* bperm;       // pair indices grouped by first Q-shell (bra order)
    const int* colf;        // compact column -> basis function (functions of every Q-shell in the block)
    int n, K, cls, mol;
    int ncol;               // number of compact columns
    int pad;
};

struct MolDesc {
    const double* P;        // nbasis x nbasis density, row-major
    const double* PJ;       // density of the Coulomb terms: P, or zeros in the exchange-only
                            // pass of an open-shell class built in two passes
    double* G;              // nbasis x nbasis accumulator (G = Gt + Gt^T afterwards)
    const double* cutmat;   // nshell x nshell max |P| per original shell pair
    int nbasis, nshell;
    double cutoff, primlim, hyb;
    double split;           // quartets with Ycut*Ycut >= split run in FP64, below in FP32
    int replicas;           // G has `replicas` copies of nbasis^2; a wave adds into one of them
    const double* P2;       // open-shell pass: the second exchange density (P_alpha - P_beta)
    double* G2;             // open-shell pass: its accumulator, replicas x nbasis^2 like G
    int natom;              // gradient: atoms (3 natom entries of grad)
    const int* shell_atom;  // gradient: original shell -> atom
    double* grad;           // gradient: 3 natom accumulator, x y z per atom
};

// One wave: bras bperm[kb..ke) of one bra block - all with the same first
// Q-shell a, Schwarz-descending - against kets [ket_start, ket_start +
// ket_count) of one ket block, one ket per lane.
struct Task {
    int bra_block, kb, ke, ket_block, ket_start, ket_count, same_block;
    int all_fp64;   // 1: this class runs every quartet in FP64 even in mixed mode
    // primitive-pair ranges of the bra and ket pairs this task sums over
    // ([0, K) unless a highly contracted class is spread over several tasks:
    // the HRR and the digestion are linear in the primitive sums, so the
    // pieces add up to the full quartet)
    int pb, pe, qb, qe;
};

// One workgroup of the LDS-tile kernels: tasks [task_begin, task_end), all
// from one bra chunk bperm[chunk_kb .. chunk_kb + nbra) of one bra block
// against one ket block.  The exchange terms of the chunk are accumulated
// in an LDS tile (rows: the functions of a and of every b_k; columns: the
// ket block's compact columns) and added to G once, at the end.
struct TileGroup {
    int task_begin, task_end;
    int chunk_kb, nbra;
};

}  // namespace qf
