// COSY map files: host-side reader and the packed device tables.
//
// Reads the forward sequential maps (format of shared/transp.f:484,
// "(1x,5g14.7,1x,6i1)") and the reconstruction maps (hms/mc_hms_recon.f:144,
// "(1x,4g16.9,1x,5i1)") exactly as transp_init / mc_*_recon do, checks them
// against the polynomial structure compiled into the kernels, and packs the
// non-zero coefficients in the order the generated device code reads them.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace simc {

struct CosyTerm {
    double c[5];
    uint8_t e[5];   // exponents of x, x', y, y', delta (forward) / x, x', y, y', x_tar (recon)
};

struct ForwardMap {
    int n_classes = 0;
    std::vector<std::vector<CosyTerm>> classes;  // TOF line dropped
    std::vector<double> length;                  // cm, !LENGTH: comments (0 when absent)
    std::vector<int> adrift;                     // pure field-free drift
    std::vector<double> driftdist;               // cm (1000 * <x|x'>)
};

struct ReconMap {
    std::vector<CosyTerm> terms;  // 4 outputs used
    int max_order = 0;
};

// Return 0 or a negative SIMC_E_* code; msg receives a description.
int read_forward_map(const std::string& path, ForwardMap& m, std::string& msg);
int read_recon_map(const std::string& path, ReconMap& m, std::string& msg);

// Pack an arm's coefficients for the generated kernels.  arm: 1 HMS, 5 SHMS.
int pack_arm(int arm, const ForwardMap& f, const ReconMap& r, std::vector<double>& packed,
             std::string& msg);

// Algorithmic FP64 operations of one pass through the polynomial of a class
// (cls = 0: recon): monomial multiplies + 2 per non-zero coefficient.
int64_t poly_flops(int arm, int cls);
int64_t poly_terms(int arm, int cls);

// Run-time polynomial tables, for maps whose term structure differs from the
// compiled one (another tune or fit of the optics).  Each class (0 = recon)
// is a list of monomial-trie nodes in depth-first order: node = (depth d,
// variable v, output mask), its monomial = the monomial of the last node of
// depth d-1 times input v, so every node costs one multiply; parents a map
// lacks are inserted without coefficients.  Coefficients of each node's
// non-zero outputs follow in output order.  One device buffer:
//   uint32 words [0, 33)   first node of class c
//   uint32 words [33, 66)  node count of class c
//   uint32 word  66        word offset of the node records, 67: double offset
//                          of the coefficients (in doubles from the start)
//   node k: word rec + 2k = d | v << 4 | mask << 8, rec + 2k + 1 = its first
//   coefficient (index into the coefficient doubles)
constexpr int kGenMaxDepth = 8;
struct GenericPoly {
    std::vector<uint32_t> words;  // the header and node records
    std::vector<double> coef;
    int64_t flops[33] = {};      // multiplies + 2 per coefficient, per class
    int64_t terms[33] = {};      // the map's terms, per class
    size_t bytes() const;        // size of the device buffer
    void pack(std::vector<uint8_t>& out) const;
};
int build_generic(const ForwardMap& f, const ReconMap& r, GenericPoly& g, std::string& msg);

}  // namespace simc
