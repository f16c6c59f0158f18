#include "mech.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace cg {

namespace {

double to_d(const std::string& s) {
    char* end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) throw std::runtime_error("bad number '" + s + "' in mechanism file");
    return v;
}

int to_i(const std::string& s) {
    char* end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str()) throw std::runtime_error("bad integer '" + s + "' in mechanism file");
    return static_cast<int>(v);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

HostMech load_mech(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open mechanism file " + path);
    HostMech m;
    std::string line;
    bool header = false, ended = false;
    int last_r = -1;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        if (line.rfind("E ", 0) == 0) {
            if (last_r >= 0) m.rxn[last_r].equation = line.substr(2);
            continue;
        }
        std::istringstream is(line);
        std::vector<std::string> t;
        std::string w;
        while (is >> w) t.push_back(w);
        if (t.empty()) continue;
        size_t p = 1;
        auto next = [&]() -> const std::string& {
            if (p >= t.size()) throw std::runtime_error("truncated record in " + path + ": " + line);
            return t[p++];
        };
        const std::string& key = t[0];
        if (key == "CGMECH") {
            header = true;
        } else if (key == "phase") {
            m.phase = next();
        } else if (key == "source") {
        } else if (key == "elements") {
            m.E = to_i(next());
            for (int e = 0; e < m.E; e++) {
                m.elements.push_back(next());
                m.aw.push_back(to_d(next()));
            }
        } else if (key == "species") {
            m.K = to_i(next());
            m.names.resize(m.K);
            m.mw.resize(m.K);
            m.comp.assign(static_cast<size_t>(m.K) * m.E, 0.0);
            m.tmin.resize(m.K); m.tmid.resize(m.K); m.tmax.resize(m.K); m.pref.resize(m.K);
            m.clow.resize(7 * m.K); m.chigh.resize(7 * m.K);
        } else if (key == "S") {
            int k = to_i(next());
            if (k < 0 || k >= m.K) throw std::runtime_error("species index out of range");
            m.names[k] = next();
            m.mw[k] = to_d(next());
            m.tmin[k] = to_d(next()); m.tmid[k] = to_d(next()); m.tmax[k] = to_d(next()); m.pref[k] = to_d(next());
            for (int j = 0; j < 7; j++) m.clow[7 * k + j] = to_d(next());
            for (int j = 0; j < 7; j++) m.chigh[7 * k + j] = to_d(next());
            for (int e = 0; e < m.E; e++) m.comp[static_cast<size_t>(k) * m.E + e] = to_d(next());
        } else if (key == "reactions") {
            m.I = to_i(next());
            m.rxn.resize(m.I);
        } else if (key == "R") {
            int i = to_i(next());
            if (i < 0 || i >= m.I) throw std::runtime_error("reaction index out of range");
            HostReaction& r = m.rxn[i];
            r.type = to_i(next());
            r.rev = to_i(next()) != 0;
            r.dn = to_d(next());
            r.A = to_d(next()); r.b = to_d(next()); r.EaR = to_d(next());
            r.A0 = to_d(next()); r.b0 = to_d(next()); r.EaR0 = to_d(next());
            for (int j = 0; j < 4; j++) r.troe[j] = to_d(next());
            int nr = to_i(next());
            for (int j = 0; j < nr; j++) {
                r.rk.push_back(to_i(next())); r.rs.push_back(to_d(next())); r.ro.push_back(to_d(next()));
            }
            int np = to_i(next());
            for (int j = 0; j < np; j++) {
                r.pk.push_back(to_i(next())); r.ps.push_back(to_d(next()));
            }
            r.tbdef = to_d(next());
            int ntb = to_i(next());
            for (int j = 0; j < ntb; j++) {
                r.tbk.push_back(to_i(next())); r.tbe.push_back(to_d(next()));
            }
            for (int k : r.rk) if (k < 0 || k >= m.K) throw std::runtime_error("reactant index out of range");
            for (int k : r.pk) if (k < 0 || k >= m.K) throw std::runtime_error("product index out of range");
            for (int k : r.tbk) if (k < 0 || k >= m.K) throw std::runtime_error("efficiency index out of range");
            if (r.type < 0 || r.type > 7) throw std::runtime_error("unknown reaction type");
            last_r = i;
        } else if (key == "PLOG") {
            if (last_r < 0 || m.rxn[last_r].type != kPlog) throw std::runtime_error("PLOG record without a PLOG reaction");
            const int ne = to_i(next());
            if (ne < 1) throw std::runtime_error("PLOG reaction without rate expressions");
            for (int e = 0; e < 4 * ne; e++) m.rxn[last_r].plog.push_back(to_d(next()));
        } else if (key == "CHEB") {
            if (last_r < 0 || m.rxn[last_r].type != kChebyshev)
                throw std::runtime_error("CHEB record without a Chebyshev reaction");
            std::vector<double>& c = m.rxn[last_r].cheb;
            for (int e = 0; e < 4; e++) c.push_back(to_d(next()));
            const int nT = to_i(next()), nP = to_i(next());
            if (nT < 1 || nP < 1) throw std::runtime_error("empty Chebyshev coefficient matrix");
            c.push_back(nT);
            c.push_back(nP);
            for (int e = 0; e < nT * nP; e++) c.push_back(to_d(next()));
        } else if (key == "CHEMACT") {
            if (last_r < 0 || !(m.rxn[last_r].type == kLindemann || m.rxn[last_r].type >= kTroe) ||
                m.rxn[last_r].type == kPlog || m.rxn[last_r].type == kChebyshev)
                throw std::runtime_error("CHEMACT record without a falloff reaction");
            m.rxn[last_r].chem_act = true;
        } else if (key == "SRI") {
            if (last_r < 0 || m.rxn[last_r].type != kSri) throw std::runtime_error("SRI record without an SRI reaction");
            for (int e = 0; e < 5; e++) m.rxn[last_r].sri[e] = to_d(next());
        } else if (key == "TSANG") {
            if (last_r < 0 || m.rxn[last_r].type != kTsang)
                throw std::runtime_error("TSANG record without a Tsang reaction");
            for (int e = 0; e < 2; e++) m.rxn[last_r].tsang[e] = to_d(next());
        } else if (key == "end") {
            ended = true;
        }
    }
    if (!header || !ended || m.K <= 0) throw std::runtime_error("not a complete .mech file: " + path);
    return m;
}

template <class T>
const T* DeviceMech::upload(const std::vector<T>& v) {
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc(mechanism table)");
    allocs_.push_back(p);
    if (!v.empty()) hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(mechanism table)");
    bytes_ += bytes;
    return static_cast<const T*>(p);
}

namespace {

// StoichManagerN::add (StoichManager.h:676-723): unit-stoichiometry sides with
// at most three molecules become repeated-species lists; everything else is
// the general (power-law) form.
bool expand_side(const std::vector<int>& k, const std::vector<double>& order,
                 const std::vector<double>& st, std::vector<int>& krep) {
    bool frac = false;
    for (size_t n = 0; n < st.size(); n++)
        if (std::fmod(st[n], 1.0) != 0.0 || st[n] != order[n]) frac = true;
    krep.clear();
    if (frac || k.size() > 3) return false;
    for (size_t n = 0; n < k.size(); n++)
        for (int i = 0; i < st[n]; i++) krep.push_back(k[n]);
    return krep.size() >= 1 && krep.size() <= 3;
}

}  // namespace

DeviceMech::DeviceMech(const HostMech& h) : h_(h) {
    const int K = h.K, I = h.I;
    // K + 2 <= 256: four state slots per lane (the dense integrators and
    // Jacobians take K <= 126, the mole-based reactors and the rates K <= 254)
    if (K > 254) throw std::runtime_error("mechanisms with more than 254 species are not supported");
    perm_.resize(I);
    std::iota(perm_.begin(), perm_.end(), 0);
    // Reactions of the costliest class first: a wavefront evaluates 64
    // consecutive reactions, so the falloff reactions share one pass (with the
    // three-body ones and the first elementary ones) instead of leaving a last
    // pass that runs the Troe code for a few lanes.  PLOG and Chebyshev
    // reactions (types 4, 5) lead: their rate constants are evaluated once per
    // state by lanes 0..63 before the reaction loop.
    // Class order: Chebyshev, PLOG, Tsang, SRI, Troe, Lindemann, three-body, elementary.
    static const int kRank[8] = {7, 6, 5, 4, 1, 0, 3, 2};   // by RxnType
    std::stable_sort(perm_.begin(), perm_.end(),
                     [&](int a, int b) { return kRank[h.rxn[a].type] < kRank[h.rxn[b].type]; });

    std::vector<double> rd(4 * static_cast<size_t>(I)), tbdef(I), fod, gend, pdd;
    std::vector<RxnI> ri(I);
    std::vector<int> tbk;
    std::vector<double> tbe;
    int n_el = 0, n_tb = 0, n_fo = 0, n_pd = 0;
    for (int p = 0; p < I; p++) {
        const HostReaction& r = h.rxn[perm_[p]];
        RxnI& x = ri[p];
        std::memset(&x, 0, sizeof x);
        for (int j = 0; j < kMaxSide; j++) { x.rk[j] = static_cast<short>(K); x.pk[j] = static_cast<short>(K); }
        x.orig = perm_[p];
        x.fo = -1;
        x.gen = -1;
        // PLOG and Chebyshev reactions are mass-action reactions (type 0 on
        // the device) whose rate constant comes from their pdd record
        // SRI and Tsang are falloff reactions of the Troe form (type 3; SRI
        // marked by bit 8, Tsang expressed through the Troe parameters).
        const int dtype = (r.type == kPlog || r.type == kChebyshev) ? kElementary
                          : (r.type == kSri || r.type == kTsang) ? kTroe : r.type;
        x.flags = (r.rev ? 1 : 0) | (dtype << 1) | (dtype != kElementary ? 32 : 0) | (r.type == kPlog ? 64 : 0) |
                  (r.type == kChebyshev ? 128 : 0) | (r.type == kSri ? 256 : 0) | (r.chem_act ? 512 : 0);
        if (r.type == kPlog || r.type == kChebyshev) n_pd++;
        if (dtype == kElementary) n_el++;
        else if (dtype == kThreeBody) n_tb++;
        else n_fo++;
        rd[4 * p + 0] = r.A; rd[4 * p + 1] = r.b; rd[4 * p + 2] = r.EaR; rd[4 * p + 3] = r.dn;
        tbdef[p] = r.tbdef;
        std::vector<int> krep;
        bool rgen = false, pgen = false;
        if (expand_side(r.rk, r.ro, r.rs, krep)) {
            x.nr = static_cast<int>(krep.size());
            for (int j = 0; j < x.nr; j++) x.rk[j] = static_cast<short>(krep[j]);
        } else {
            rgen = true;
            if (r.rk.size() > kMaxSide) throw std::runtime_error("reaction '" + r.equation + "' has too many reactants");
            x.nr = static_cast<int>(r.rk.size());
            for (int j = 0; j < x.nr; j++) x.rk[j] = static_cast<short>(r.rk[j]);
        }
        // products are kept for every reaction (the Jacobian needs them); only
        // reversible reactions multiply them into the reverse rate
        {
            if (expand_side(r.pk, r.ps, r.ps, krep)) {
                x.np = static_cast<int>(krep.size());
                for (int j = 0; j < x.np; j++) x.pk[j] = static_cast<short>(krep[j]);
            } else {
                pgen = true;
                if (r.pk.size() > kMaxSide) throw std::runtime_error("reaction '" + r.equation + "' has too many products");
                x.np = static_cast<int>(r.pk.size());
                for (int j = 0; j < x.np; j++) x.pk[j] = static_cast<short>(r.pk[j]);
            }
        }
        if (dtype == kElementary && !(x.flags & (64 | 128)) && !rgen && !pgen) x.flags |= 1024;
        if (rgen || pgen) {
            x.flags |= (rgen ? 8 : 0) | (pgen ? 16 : 0);
            x.gen = static_cast<int>(gend.size() / 12);
            double g[12] = {0};
            if (rgen)
                for (int j = 0; j < x.nr; j++) { g[j] = r.ro[j]; g[4 + j] = r.rs[j]; }
            if (pgen)
                for (int j = 0; j < x.np; j++) g[8 + j] = r.ps[j];
            gend.insert(gend.end(), g, g + 12);
        }
        if (dtype >= kLindemann) {
            x.fo = static_cast<int>(fod.size() / 8);
            double fo[8] = {r.A0, r.b0, r.EaR0, r.troe[0], r.troe[1], r.troe[2], r.troe[3], 0.0};
            if (r.type == kTsang) {
                // Fcent = A + B T (TsangRate::updateTemp, Falloff.cpp:480-484) as
                // (1 - A) e^(-T inf) + A e^(-T 0) + B T
                fo[3] = r.tsang[0]; fo[4] = INFINITY; fo[5] = 0.0; fo[6] = 0.0; fo[7] = r.tsang[1];
            } else if (r.type == kSri) {
                for (int j = 0; j < 5; j++) fo[3 + j] = r.sri[j];
            }
            fod.insert(fod.end(), fo, fo + 8);
        }
        if (r.type == kPlog) {
            // PlogRate::setRates (PlogRate.cpp:92-123): groups of equal
            // pressure, plus copies of the first and last group at
            // ln P = -1000 and +1000.  Record: ng, ln P[ng], first[ng],
            // end[ng] (entry indices), then A, b, Ea/R per entry.
            const int ne = static_cast<int>(r.plog.size() / 4);
            std::vector<double> lp, g0, g1;
            for (int e = 0; e < ne; e++) {
                const double l = std::log(r.plog[4 * e]);
                if (lp.empty() || lp.back() != l) { lp.push_back(l); g0.push_back(e); g1.push_back(e + 1); }
                else g1.back() = e + 1;
            }
            lp.insert(lp.begin(), -1000.0); g0.insert(g0.begin(), g0.front()); g1.insert(g1.begin(), g1.front());
            lp.push_back(1000.0); g0.push_back(g0.back()); g1.push_back(g1.back());
            x.fo = static_cast<int>(pdd.size());
            pdd.push_back(static_cast<double>(lp.size()));
            pdd.insert(pdd.end(), lp.begin(), lp.end());
            pdd.insert(pdd.end(), g0.begin(), g0.end());
            pdd.insert(pdd.end(), g1.begin(), g1.end());
            for (int e = 0; e < ne; e++)
                for (int j = 1; j < 4; j++) pdd.push_back(r.plog[4 * e + j]);
        } else if (r.type == kChebyshev) {
            // ChebyshevRate::setLimits (ChebyshevRate.cpp:96-112).  Record:
            // TrNum, TrDen, PrNum, PrDen, nT, nP, coefficients (row-major).
            const double Tmin = r.cheb[0], Tmax = r.cheb[1];
            const double lPmin = std::log10(r.cheb[2]), lPmax = std::log10(r.cheb[3]);
            const double TminInv = 1.0 / Tmin, TmaxInv = 1.0 / Tmax;
            x.fo = static_cast<int>(pdd.size());
            pdd.push_back(-TminInv - TmaxInv);
            pdd.push_back(1.0 / (TmaxInv - TminInv));
            pdd.push_back(-lPmin - lPmax);
            pdd.push_back(1.0 / (lPmax - lPmin));
            pdd.insert(pdd.end(), r.cheb.begin() + 4, r.cheb.end());
        }
        if (dtype != kElementary) {
            if (r.tbk.size() > kMaxTB) throw std::runtime_error("too many third-body efficiencies");
            x.tbo = static_cast<int>(tbk.size());
            x.tbn = static_cast<int>(r.tbk.size());
            tbk.insert(tbk.end(), r.tbk.begin(), r.tbk.end());
            tbe.insert(tbe.end(), r.tbe.begin(), r.tbe.end());
        }
    }

    std::vector<double> nasa(16 * static_cast<size_t>(K), 0.0), rmw(K);
    for (int k = 0; k < K; k++) {
        nasa[16 * k] = h.tmid[k];
        for (int j = 0; j < 7; j++) { nasa[16 * k + 1 + j] = h.clow[7 * k + j]; nasa[16 * k + 8 + j] = h.chigh[7 * k + j]; }
        rmw[k] = 1.0 / h.mw[k];
    }

    d_.K = K; d_.I = I; d_.N = K + 2; d_.E = h.E;
    d_.n_el = n_el; d_.n_tb = n_tb; d_.n_fo = n_fo;
    d_.n_pd = n_pd;
    if (n_pd > 64) throw std::runtime_error("mechanisms with more than 64 PLOG and Chebyshev reactions are not supported");
    d_.pref = h.pref.empty() ? 101325.0 : h.pref[0];
    d_.mw = upload(h.mw);
    d_.rmw = upload(rmw);
    d_.nasa = upload(nasa);
    d_.rd = upload(rd);
    d_.tbdef = upload(tbdef);
    d_.ri = upload(ri);
    d_.fod = upload(fod);
    d_.gend = upload(gend);
    d_.pdd = upload(pdd);
    // kTbPad entries of padding (species K, whose concentration slot holds 1,
    // efficiency 0) after the last list: the first kTbPad efficiencies of any
    // reaction are loaded without a test against its list length
    tbk.insert(tbk.end(), kTbPad, K);
    tbe.insert(tbe.end(), kTbPad, 0.0);
    d_.tbk = upload(tbk);
    d_.tbe = upload(tbe);
    d_.comp = upload(h.comp);
}

DeviceMech::~DeviceMech() {
    for (void* p : allocs_) (void)hipFree(p);
}

}  // namespace cg
