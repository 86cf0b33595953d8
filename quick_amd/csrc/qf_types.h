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

enum PrimField { PF_ZETA = 0, PF_OOZ2, PF_PX, PF_PY, PF_PZ, PF_KC, PF_CUT, PF_N };

struct PairInfo {
    int qa, qb;      // Q-shell ids (an SP shell is split into an s and a p Q-shell)
    int fa, fb;      // first basis function of the two Q-shells
    int sa, sb;      // original shell ids (density-screen matrix index)
    int qfs;         // Qfinal(orig a) + Qfinal(orig b): QUICK's NABCD contribution
    int id;          // unique pair id within the molecule
    double ycut;     // Schwarz bound of the original shell pair
    double A[3];     // centre of Q-shell a
    double AB[3];    // A - B
    double pad;
};

struct PairBlock {
    const PairInfo* info;   // n entries
    const double* prim;     // PF_N * K * n
    const int* bperm;       // pair indices grouped by first Q-shell (bra order)
    int n, K, cls, mol;
};

struct MolDesc {
    const double* P;        // nbasis x nbasis density, row-major
    double* G;              // nbasis x nbasis accumulator (G = Gt + Gt^T afterwards)
    const double* cutmat;   // nshell x nshell max |P| per original shell pair
    int nbasis, nshell;
    double cutoff, primlim, hyb;
    double split;           // quartets with Ycut*Ycut >= split run in FP64, below in FP32
    int replicas;           // G has `replicas` copies of nbasis^2; a wave adds into one of them
    int pad;
};

// One wave: bras bperm[kb..ke) of one bra block - all with the same first
// Q-shell a, Schwarz-descending - against kets [ket_start, ket_start +
// ket_count) of one ket block, one ket per lane.
struct Task {
    int bra_block, kb, ke, ket_block, ket_start, ket_count, same_block;
    int all_fp64;   // 1: this class runs every quartet in FP64 even in mixed mode
};

}  // namespace qf
