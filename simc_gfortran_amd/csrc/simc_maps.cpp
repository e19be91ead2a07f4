// COSY map reader and packer (host).  See simc_maps.h.
#include "simc_maps.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <map>

#include "../../include/simc_gpu.h"
#include "cosy_gen.host.inc"

namespace simc {
namespace {

constexpr int kMaxClass = 41;    // spectrometers.inc:5
constexpr int kMaxTerms = 1000;  // shared/transp.f:81

// One formatted record into a fixed 132-column image (CHARACTER*132 read with
// format (a): longer records are cut, shorter ones blank-padded).
bool next_record(std::ifstream& f, std::string& rec)
{
    std::string line;
    if (!std::getline(f, line)) return false;
    if (line.size() > 132) line.resize(132);
    rec.assign(132, ' ');
    rec.replace(0, line.size(), line);
    return true;
}

// Real input editing of a w-column field, BLANK='NULL'.  Without a decimal
// point the last d digits of the mantissa are the fraction (Fortran implied
// decimal point); exponents may be written E, D, Q or as a bare signed number.
bool parse_real(const std::string& rec, size_t pos, int w, int d, double& out)
{
    std::string s;
    for (int i = 0; i < w; ++i) {
        char ch = rec[pos + i];
        if (ch != ' ' && ch != '\t') s.push_back(ch);
    }
    if (s.empty()) { out = 0.0; return true; }
    size_t i = 0;
    bool neg = false;
    if (s[i] == '+' || s[i] == '-') neg = s[i++] == '-';
    std::string mant;
    bool point = false;
    int frac = 0;
    for (; i < s.size(); ++i) {
        if (std::isdigit((unsigned char)s[i])) { mant.push_back(s[i]); if (point) ++frac; }
        else if (s[i] == '.' && !point) point = true;
        else break;
    }
    if (mant.empty()) return false;
    long ex = 0;
    if (i < s.size()) {
        char c = s[i];
        if (c == 'E' || c == 'e' || c == 'D' || c == 'd' || c == 'Q' || c == 'q') ++i;
        bool eneg = false;
        if (i < s.size() && (s[i] == '+' || s[i] == '-')) eneg = s[i++] == '-';
        if (i >= s.size()) return false;
        for (; i < s.size(); ++i) {
            if (!std::isdigit((unsigned char)s[i])) return false;
            ex = ex * 10 + (s[i] - '0');
        }
        if (eneg) ex = -ex;
    }
    long scale = point ? -frac : -d;
    std::string canon = (neg ? "-" : "") + mant + "e" + std::to_string(ex + scale);
    out = std::strtod(canon.c_str(), nullptr);  // correctly rounded
    return true;
}

bool parse_i1(char c, int& v)
{
    if (c == ' ') { v = 0; return true; }
    if (c >= '0' && c <= '9') { v = c - '0'; return true; }
    return false;
}

double list_real(const std::string& rec, size_t pos)
{
    while (pos < rec.size() && (rec[pos] == ' ' || rec[pos] == ',')) ++pos;
    std::string tok;
    while (pos < rec.size() && rec[pos] != ' ' && rec[pos] != ',' && rec[pos] != '/') {
        char c = rec[pos++];
        tok.push_back(c == 'd' || c == 'D' ? 'e' : c);
    }
    return std::strtod(tok.c_str(), nullptr);
}

bool blank(const std::string& rec)
{
    for (char c : rec)
        if (c != ' ') return false;
    return true;
}

}  // namespace

// transp_init, shared/transp.f:318-474
int read_forward_map(const std::string& path, ForwardMap& m, std::string& msg)
{
    const double coeff_min = 1.0e-14;  // transp.f:66
    std::ifstream f(path);
    if (!f) { msg = "cannot open forward map " + path; return SIMC_E_IO; }
    m = ForwardMap();
    m.classes.assign(kMaxClass, {});
    m.length.assign(kMaxClass, 0.0);
    m.adrift.assign(kMaxClass, 1);
    m.driftdist.assign(kMaxClass, 0.0);
    std::string rec(132, ' ');
    rec[0] = '!';
    while (rec[0] == '!') {  // header (:325-334)
        if (!next_record(f, rec)) { msg = "forward map has no data: " + path; return SIMC_E_MAPFMT; }
        if (rec.compare(0, 8, "!LENGTH:") == 0) m.length[0] = 100. * list_real(rec, 9);
    }
    int n_classes = 0;
    for (;;) {
        int kk = n_classes;
        if (kk >= kMaxClass) { msg = "too many transformations in " + path; return SIMC_E_TOOMANY; }
        auto& cls = m.classes[kk];
        while (rec.compare(0, 4, " ---") != 0) {
            CosyTerm t{};
            int dig[6];
            for (int j = 0; j < 5; ++j)
                if (!parse_real(rec, 1 + 14 * j, 14, 7, t.c[j])) {
                    msg = "bad coefficient field in " + path;
                    return SIMC_E_MAPFMT;
                }
            for (int j = 0; j < 6; ++j)
                if (!parse_i1(rec[72 + j], dig[j])) { msg = "bad exponent field in " + path; return SIMC_E_MAPFMT; }
            t.e[0] = dig[0]; t.e[1] = dig[1]; t.e[2] = dig[2]; t.e[3] = dig[3]; t.e[4] = dig[5];
            const int tof = dig[4];
            if (tof != 0) {  // :384-391
                if (t.c[0] != 0 || t.c[1] != 0 || t.c[2] != 0 || t.c[3] != 0) {
                    msg = "non-zero TOF terms in " + path;
                    return SIMC_E_TOF;
                }
            } else {
                if ((int)cls.size() >= kMaxTerms) { msg = "too many COSY terms in " + path; return SIMC_E_TOOMANY; }
                cls.push_back(t);
                if (m.adrift[kk]) {  // drift test, :399-438
                    const int e1 = t.e[0], e2 = t.e[1], e3 = t.e[2], e4 = t.e[3];
                    const double c1 = t.c[0], c2 = t.c[1], c3 = t.c[2], c4 = t.c[3];
                    const double csum = std::fabs(c1) + std::fabs(c2) + std::fabs(c3) + std::fabs(c4);
                    const int order = e1 + e2 + e3 + e4;
                    auto off = [&](double v, double want) { if (std::fabs(v - want) > coeff_min) m.adrift[kk] = 0; };
                    if (order == 1) {
                        if (e1 == 1) { off(c1, 1.e0); off(c2, 0.e0); off(c3, 0.e0); off(c4, 0.e0); }
                        else if (e2 == 1) {
                            m.driftdist[kk] = 1000.e0 * c1;
                            off(c2, 1.e0); off(c3, 0.e0); off(c4, 0.e0);
                        } else if (e3 == 1) { off(c1, 0.e0); off(c2, 0.e0); off(c3, 1.e0); off(c4, 0.e0); }
                        else if (e4 == 1) {
                            off(c1, 0.e0); off(c2, 0.e0);
                            if (std::fabs(m.driftdist[kk] - 1000.e0 * c3) > coeff_min) m.adrift[kk] = 0;
                            off(c4, 1.e0);
                        }
                    } else if (std::fabs(csum) > coeff_min) {
                        m.adrift[kk] = 0;
                    }
                }
            }
            if (!next_record(f, rec)) { msg = "forward map ends inside a class: " + path; return SIMC_E_MAPFMT; }
        }
        n_classes = kk + 1;
        bool eof = false;
        for (;;) {  // :461-468
            if (!next_record(f, rec)) { eof = true; break; }
            if (rec.compare(0, 8, "!LENGTH:") == 0 && kk + 1 < kMaxClass) m.length[kk + 1] = 100. * list_real(rec, 9);
            if (rec[0] == '!' || rec.compare(0, 4, " ---") == 0 || blank(rec)) continue;
            break;
        }
        if (eof) break;
    }
    m.n_classes = n_classes;
    m.classes.resize(n_classes);
    return SIMC_OK;
}

// hms/mc_hms_recon.f:70-102 (shms/mc_shms_recon.f:106-128 is the same reader)
int read_recon_map(const std::string& path, ReconMap& m, std::string& msg)
{
    std::ifstream f(path);
    if (!f) { msg = "cannot open recon map " + path; return SIMC_E_IO; }
    m = ReconMap();
    std::string rec(132, ' ');
    rec[0] = '!';
    while (rec[0] == '!')
        if (!next_record(f, rec)) { msg = "recon map has no data: " + path; return SIMC_E_MAPFMT; }
    while (rec.compare(0, 4, " ---") != 0) {
        if ((int)m.terms.size() >= kMaxTerms) { msg = "too many COSY terms in " + path; return SIMC_E_TOOMANY; }
        CosyTerm t{};
        for (int j = 0; j < 4; ++j)
            if (!parse_real(rec, 1 + 16 * j, 16, 9, t.c[j])) { msg = "bad coefficient field in " + path; return SIMC_E_MAPFMT; }
        int order = 0;
        for (int j = 0; j < 5; ++j) {
            int v;
            if (!parse_i1(rec[66 + j], v)) { msg = "bad exponent field in " + path; return SIMC_E_MAPFMT; }
            t.e[j] = (uint8_t)v;
            order += v;
        }
        m.terms.push_back(t);
        if (!next_record(f, rec)) { msg = "recon map ends without ' ---': " + path; return SIMC_E_MAPFMT; }
        if (order > m.max_order) m.max_order = order;
    }
    return SIMC_OK;
}

namespace {
const simc_gen::PolyInfo* polys(int arm, int& n, int& ncoef)
{
    if (arm == SIMC_ARM_HMS) { n = simc_gen::n_polys_hms; ncoef = simc_gen::n_coef_hms; return simc_gen::polys_hms; }
    if (arm == SIMC_ARM_SOS) { n = simc_gen::n_polys_sos; ncoef = simc_gen::n_coef_sos; return simc_gen::polys_sos; }
    if (arm == SIMC_ARM_HRSR || arm == SIMC_ARM_HRSL) {
        n = simc_gen::n_polys_hrs;
        ncoef = simc_gen::n_coef_hrs;
        return simc_gen::polys_hrs;
    }
    n = simc_gen::n_polys_shms;
    ncoef = simc_gen::n_coef_shms;
    return simc_gen::polys_shms;
}
}  // namespace

int pack_arm(int arm, const ForwardMap& f, const ReconMap& r, std::vector<double>& packed,
             std::string& msg)
{
    int n, ncoef;
    const simc_gen::PolyInfo* P = polys(arm, n, ncoef);
    packed.assign(ncoef, 0.0);
    for (int i = 0; i < n; ++i) {
        const simc_gen::PolyInfo& p = P[i];
        const std::vector<CosyTerm>* terms;
        if (p.cls == 0) terms = &r.terms;
        else {
            if (p.cls > f.n_classes) { msg = "forward map lacks class " + std::to_string(p.cls); return SIMC_E_NCLASS; }
            terms = &f.classes[p.cls - 1];
        }
        const char* what = p.cls ? "forward class " : "recon map";
        if ((int)terms->size() != p.n_terms) {
            msg = std::string(what) + std::to_string(p.cls) + ": " + std::to_string(terms->size()) +
                  " terms, kernel compiled for " + std::to_string(p.n_terms);
            return SIMC_E_MAPSTRUCT;
        }
        std::vector<uint8_t> used(p.n_terms * 5, 0);
        for (int t = 0; t < p.n_terms; ++t)
            for (int v = 0; v < 5; ++v)
                if ((*terms)[t].e[v] != p.expon[t][v]) {
                    msg = std::string(what) + std::to_string(p.cls) + ": term " + std::to_string(t + 1) +
                          " exponents differ from the compiled structure";
                    return SIMC_E_MAPSTRUCT;
                }
        for (int k = 0; k < p.n_coef; ++k) {
            const int t = p.pack[k][0], o = p.pack[k][1];
            packed[p.coef_offset + k] = (*terms)[t].c[o];
            used[t * 5 + o] = 1;
        }
        for (int t = 0; t < p.n_terms; ++t)
            for (int o = 0; o < p.nout; ++o)
                if ((*terms)[t].c[o] != 0.0 && !used[t * 5 + o]) {
                    msg = std::string(what) + std::to_string(p.cls) + ": coefficient (" + std::to_string(t + 1) +
                          "," + std::to_string(o + 1) + ") is non-zero where the compiled structure has none";
                    return SIMC_E_MAPSTRUCT;
                }
    }
    return SIMC_OK;
}

namespace {

// One class's monomial trie: the terms as sorted variable sequences (x0^2 x3
// -> 0 0 3), every prefix present, in lexicographic (= depth-first) order.
int add_generic_class(const std::vector<CosyTerm>& terms, int nout, GenericPoly& g, int cls, std::string& msg)
{
    std::map<std::vector<uint8_t>, std::vector<int>> nodes;  // sequence -> terms with that monomial
    nodes[{}];
    for (int t = 0; t < (int)terms.size(); ++t) {
        std::vector<uint8_t> seq;
        for (int v = 0; v < 5; ++v)
            for (int k = 0; k < terms[t].e[v]; ++k) seq.push_back((uint8_t)v);
        if ((int)seq.size() > kGenMaxDepth) {
            msg = "class " + std::to_string(cls) + ": term of order " + std::to_string(seq.size()) +
                  " beyond the run-time polynomial's " + std::to_string(kGenMaxDepth);
            return SIMC_E_MAPSTRUCT;
        }
        nodes[seq].push_back(t);
        for (size_t d = 1; d < seq.size(); ++d) nodes[std::vector<uint8_t>(seq.begin(), seq.begin() + d)];
    }
    g.words[cls] = (uint32_t)(g.words.size() - 68) / 2;
    g.words[33 + cls] = (uint32_t)nodes.size();
    int64_t flops = 0;
    for (const auto& [seq, ts] : nodes) {
        double c[5] = {0, 0, 0, 0, 0};
        for (int t : ts)  // a monomial listed twice sums its coefficients
            for (int o = 0; o < nout; ++o) c[o] += terms[t].c[o];
        uint32_t mask = 0;
        for (int o = 0; o < nout; ++o)
            if (c[o] != 0.0) mask |= 1u << o;
        const uint32_t d = (uint32_t)seq.size(), v = d ? seq.back() : 0u;
        g.words.push_back(d | v << 4 | mask << 8);
        g.words.push_back((uint32_t)g.coef.size());
        for (int o = 0; o < nout; ++o)
            if (mask & (1u << o)) g.coef.push_back(c[o]);
        flops += (d > 1 ? 1 : 0) + 2 * __builtin_popcount(mask);
    }
    g.flops[cls] = flops;
    g.terms[cls] = (int64_t)terms.size();
    return SIMC_OK;
}

}  // namespace

int build_generic(const ForwardMap& f, const ReconMap& r, GenericPoly& g, std::string& msg)
{
    g.words.assign(68, 0u);
    g.coef.clear();
    int rc = add_generic_class(r.terms, 4, g, 0, msg);
    for (int c = 1; !rc && c <= f.n_classes && c <= 32; ++c) rc = add_generic_class(f.classes[c - 1], 5, g, c, msg);
    if (rc) return rc;
    g.words[66] = 68;
    // coefficients start at the next 8-byte boundary after the words
    g.words[67] = (uint32_t)((g.words.size() * 4 + 7) / 8);
    return SIMC_OK;
}

size_t GenericPoly::bytes() const { return (size_t)words[67] * 8 + coef.size() * 8; }

void GenericPoly::pack(std::vector<uint8_t>& out) const
{
    out.assign(bytes(), 0);
    std::memcpy(out.data(), words.data(), words.size() * 4);
    std::memcpy(out.data() + (size_t)words[67] * 8, coef.data(), coef.size() * 8);
}

int64_t poly_flops(int arm, int cls)
{
    int n, ncoef;
    const simc_gen::PolyInfo* P = polys(arm, n, ncoef);
    for (int i = 0; i < n; ++i)
        if (P[i].cls == cls) return (int64_t)P[i].n_mul + 2 * (int64_t)P[i].n_coef;
    return 0;
}

int64_t poly_terms(int arm, int cls)
{
    int n, ncoef;
    const simc_gen::PolyInfo* P = polys(arm, n, ncoef);
    for (int i = 0; i < n; ++i)
        if (P[i].cls == cls) return P[i].n_terms;
    return 0;
}

}  // namespace simc
