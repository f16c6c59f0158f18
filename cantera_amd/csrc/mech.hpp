// Host-side mechanism: parse the compiled .mech tables and build the device
// tables used by every kernel.  Replaces the reference's mechanism assembly
// (newSolution -> BulkKinetics::addReaction; src/kinetics/BulkKinetics.cpp:16-84,
// src/kinetics/Kinetics.cpp:709-800) for the ideal-gas bulk path.
#pragma once
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

namespace cg {

constexpr int kMaxSide = 4;   // species entries per reaction side on the device
constexpr int kMaxTB = 32;    // explicit efficiencies per reaction
constexpr int kTbPad = 8;     // padding entries after the efficiency table

enum RxnType { kElementary = 0, kThreeBody = 1, kLindemann = 2, kTroe = 3, kPlog = 4, kChebyshev = 5, kSri = 6, kTsang = 7 };

struct HostReaction {
    int type = 0;
    bool rev = false;
    double dn = 0, A = 0, b = 0, EaR = 0, A0 = 0, b0 = 0, EaR0 = 0;
    double troe[4] = {0, 0, 0, 0};   // a, 1/T3, 1/T1, T2
    double sri[5] = {0, 0, 0, 1, 0}; // kSri: a, b, c, d, e
    double tsang[2] = {0, 0};        // kTsang: A, B
    bool chem_act = false;           // chemically activated falloff: k = k0 F / (1 + Pr)
    std::vector<int> rk;             // reactant species (name order)
    std::vector<double> rs, ro;      // stoich, order
    std::vector<int> pk;
    std::vector<double> ps;
    double tbdef = 0;
    std::vector<int> tbk;
    std::vector<double> tbe;         // efficiency - default
    std::vector<double> plog;        // kPlog: (P [Pa], A, b, Ea/R) per entry, sorted by P
    std::vector<double> cheb;        // kChebyshev: Tmin, Tmax, Pmin, Pmax, nT, nP, nT x nP coefficients
    std::string equation;
};

struct HostMech {
    std::string phase;
    int K = 0, I = 0, E = 0;
    std::vector<std::string> elements, names;
    std::vector<double> aw, mw, comp;              // comp: K x E
    std::vector<double> tmin, tmid, tmax, pref;
    std::vector<double> clow, chigh;               // K x 7
    std::vector<HostReaction> rxn;
};

// Parse a .mech file written by cantera_amd/mechc.py.  Throws std::runtime_error.
HostMech load_mech(const std::string& path);

// Per-reaction device record (integer part).  Species lists follow the
// reference StoichManagerN convention: unit-stoichiometry sides with at most
// three molecules are expanded into repeated species ("kRep", C1/C2/C3); other
// sides are "general" (C_AnyN: species, order, stoich).  Unused entries of
// rk/pk hold K, a sentinel species whose concentration is 1 and chemical
// potential 0 in the kernels' scratch, so a plain mass-action reaction can be
// evaluated without tests on nr and np.
struct RxnI {
    int flags;      // bit0 reversible, bits1-2 type (0..3), bit3 reactant side general, bit4 product side general,
                    // bit5 third body, bit6 PLOG, bit7 Chebyshev (those two with type 0),
                    // bit8 SRI (with type 3), bit9 chemically activated (types 2, 3),
                    // bit10 plain mass action (type 0, no PLOG/Chebyshev, both sides kRep)
    int nr, np;     // entries in rk/pk
    int tbo, tbn;   // third-body efficiency list offset/count
    int fo;         // falloff record index, or PLOG / Chebyshev record offset in pdd (-1 if none)
    int gen;        // general-order record index (-1 if none)
    int orig;       // reaction index in mechanism order
    short rk[kMaxSide];
    short pk[kMaxSide];
};
static_assert(sizeof(RxnI) == 48, "RxnI layout");

// Device view of a mechanism (all pointers are device pointers).
struct DevMech {
    int K, I, N, E;          // species, reactions, state size (K+2), elements
    int n_el, n_tb, n_fo;    // reactions of each class, in permuted order:
                             // [Chebyshev | PLOG | Troe | Lindemann | three-body | elementary]
    int n_pd;                // PLOG + Chebyshev reactions (the first n_pd, at most 64)
    double pref;
    const double* mw;        // K
    const double* rmw;       // K
    const double* nasa;      // K x 16: tmid, low[7], high[7], pad
    const double* rd;        // I x 4: A, b, EaR, dn
    const double* tbdef;     // I
    const RxnI* ri;          // I
    const double* fod;       // n_fo x 8: A0, b0, EaR0, then Troe a, 1/T3, 1/T1, T2, B (Tsang: a = A, 1/T3 = inf,
                             // 1/T1 = 0, T2 = 0, B) or SRI a, b, c, d, e
    const double* gend;      // n_gen x 12: reactant order[4], reactant stoich[4], product stoich[4]
    const int* tbk;          // efficiency species
    const double* tbe;       // efficiency - default
    const double* comp;      // K x E element composition
    const double* pdd;       // PLOG and Chebyshev records (see DeviceMech)
    int const_volume;        // reactor model of an integrator launch: 0 IdealGasConstPressureReactor
                             // (the P argument is the pressure), 1 IdealGasReactor (it is the volume)
};

// Owns device memory for a DevMech.
class DeviceMech {
public:
    explicit DeviceMech(const HostMech& h);
    ~DeviceMech();
    DeviceMech(const DeviceMech&) = delete;
    DeviceMech& operator=(const DeviceMech&) = delete;
    const DevMech& view() const { return d_; }
    const HostMech& host() const { return h_; }
    const std::vector<int>& perm() const { return perm_; }   // permuted index -> mechanism index
    size_t table_bytes() const { return bytes_; }

private:
    HostMech h_;
    DevMech d_{};
    std::vector<void*> allocs_;
    std::vector<int> perm_;
    size_t bytes_ = 0;
    template <class T> const T* upload(const std::vector<T>& v);
};

}  // namespace cg
