// ref_harness.cpp -- drives the reference's own C++ hot path (compiled from
// its sources in place; see oracle/Makefile) to write golden vectors.
//
// TEST INFRASTRUCTURE ONLY.  This file is original harness code: it includes
// the reference headers, instantiates the reference types named by the Delta-U
// path and prints what they compute.  Nothing here is shipped; the binary lands
// in oracle/_ref/ and is run by tests/golden/make_golden.py in the build
// container to produce the committed fixtures.
//
// Usage:
//   ref_harness rng <seed> <n>
//   ref_harness energy <kind> <config.json> <statefile>
//   ref_harness mc <kind> <config.json> <statefile> <ncalls>
//   ref_harness mctime <kind> <config.json> <statefile> <ncalls>   (CPU baseline)
//   ref_harness mc <kind>_npt <config.json> <statefile> <ncalls> [<ncalls_at>]
//   ref_harness mcprop <kind>_npt <config.json> <statefile> <ncalls>  (one Propagator, all moves)
// kind: dhhs | coulhs | lj | dh | ewaldlj | dhhs_cut | dhhs_vec | ljcut | dhhs_npt |
//       ewaldlj_npt
// Output: one JSON object on stdout, doubles printed with %.17g.

#include <faunus/faunus.h>
#include <faunus/ewald.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>

using namespace Faunus;
using namespace Faunus::Potential;

typedef Geometry::Cuboid Tgeo;
typedef Space<Tgeo, PointParticle> Tspace;

static void jnum(const char *key, double v, bool comma = true)
{
    if (std::isinf(v)) std::printf("\"%s\": \"%s\"", key, v > 0 ? "inf" : "-inf");
    else if (std::isnan(v)) std::printf("\"%s\": \"nan\"", key);
    else std::printf("\"%s\": %.17g", key, v);
    if (comma) std::printf(",\n");
}

static void jvec(const char *key, const std::vector<double> &v, bool comma = true)
{
    std::printf("\"%s\": [", key);
    for (size_t i = 0; i < v.size(); ++i) {
        if (std::isinf(v[i])) std::printf("\"%s\"", v[i] > 0 ? "inf" : "-inf");
        else if (std::isnan(v[i])) std::printf("\"nan\"");
        else std::printf("%.17g", v[i]);
        if (i + 1 < v.size()) std::printf(",");
    }
    std::printf("]");
    if (comma) std::printf(",\n");
}

static void jivec(const char *key, const std::vector<int> &v, bool comma = true)
{
    std::printf("\"%s\": [", key);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%d%s", v[i], i + 1 < v.size() ? "," : "");
    std::printf("]");
    if (comma) std::printf(",\n");
}

// --------------------------------------------------------------------------
// Recording moves: subclasses of the reference moves that log the energy
// change each trial returns and whether it was accepted.
// --------------------------------------------------------------------------
struct Recorder {
    std::vector<double> du;
    std::vector<int> acc;
    std::vector<int> who;
};

template<class Tspace>
class RecAT : public Move::AtomicTranslation<Tspace> {
    typedef Move::AtomicTranslation<Tspace> base;
public:
    Recorder *rec;
    RecAT(Energy::Energybase<Tspace> &e, Tspace &s, Tmjson &j, Recorder *r) : base(e, s, j), rec(r) {}
    static RandomTwister<> &internal() { return Move::Movebase<Tspace>::_slump(); }
protected:
    double _energyChange() override {
        double du = base::_energyChange();
        rec->du.push_back(du);
        rec->who.push_back(this->iparticle);
        return du;
    }
    void _acceptMove() override { rec->acc.push_back(1); base::_acceptMove(); }
    void _rejectMove() override { rec->acc.push_back(0); base::_rejectMove(); }
};

template<class Tspace>
class RecTR : public Move::TranslateRotate<Tspace> {
    typedef Move::TranslateRotate<Tspace> base;
public:
    Recorder *rec;
    RecTR(Energy::Energybase<Tspace> &e, Tspace &s, Tmjson &j, Recorder *r) : base(e, s, j), rec(r) {}
protected:
    double _energyChange() override {
        double du = base::_energyChange();
        rec->du.push_back(du);
        rec->who.push_back(this->spc->findIndex(this->igroup));
        return du;
    }
    void _acceptMove() override { rec->acc.push_back(1); base::_acceptMove(); }
    void _rejectMove() override { rec->acc.push_back(0); base::_rejectMove(); }
};


template<class Tspace>
class RecIso : public Move::Isobaric<Tspace> {
    typedef Move::Isobaric<Tspace> base;
public:
    Recorder *rec;
    template<class Tenergy>
    RecIso(Tenergy &e, Tspace &s, Tmjson &j, Recorder *r) : base(e, s, j), rec(r) {}
protected:
    double _energyChange() override {
        double du = base::_energyChange();
        rec->du.push_back(du);
        rec->who.push_back(-1);
        return du;
    }
    void _acceptMove() override { rec->acc.push_back(1); base::_acceptMove(); }
    void _rejectMove() override { rec->acc.push_back(0); base::_rejectMove(); }
};


// Atom-type table and Bjerrum length as the reference derived them from the
// JSON input (species.h:279-417 unit conversions, physconst.h:91-94).
static void print_params(Tmjson &in)
{
    std::vector<double> q, r, sig, eps, dp, mw;
    for (auto &a : atom) {
        q.push_back(a.charge); r.push_back(a.radius); sig.push_back(a.sigma);
        eps.push_back(a.eps); dp.push_back(a.dp); mw.push_back(a.mw);
    }
    jvec("type_charge", q); jvec("type_radius", r); jvec("type_sigma", sig);
    jvec("type_eps", eps); jvec("type_dp", dp); jvec("type_mw", mw);
    double epsr = in["energy"]["nonbonded"].value("epsr", 80.0);
    jnum("lB", pc::lB(epsr));
}

// Number of k-vectors in use as the energy's own info() reports it
// (NonbondedEwald::_info, ewald.h:417-470: "Wavefunctions  K (represented)");
// -1 for energies without k-space.
template<class Tenergy>
static int info_kvectors(Tenergy &pot)
{
    std::string s = pot.info();
    size_t k = s.find("Wavefunctions");
    if (k == std::string::npos) return -1;
    k = s.find_first_of("0123456789", k);
    return k == std::string::npos ? -1 : std::atoi(s.c_str() + k);
}

// Every Energybase entry point of `pot` on the loaded space (prints the
// JSON fields after the parameters).
template<class Tenergy>
static void print_energies(Tspace &spc, Tenergy &pot)
{
    pot.setSpace(spc);
    int n = (int)spc.p.size();
    double usys = Energy::systemEnergy(spc, pot, spc.p);
    jnum("system_energy", usys);
    jnum("ewald_nk", info_kvectors(pot));
    jnum("external", pot.external(spc.p));
    std::vector<double> gext, gint;
    for (auto g : spc.groupList()) {
        gext.push_back(pot.g_external(spc.p, *g));
        gint.push_back(pot.g_internal(spc.p, *g));
    }
    jvec("g_external", gext);
    jvec("g_internal", gint);
    std::vector<double> g2g;
    std::vector<int> g2g_ij;
    int ng = (int)spc.groupList().size();
    for (int a = 0; a < ng; ++a)
        for (int b = 0; b < ng; ++b) {
            g2g.push_back(pot.g2g(spc.p, *spc.groupList()[a], *spc.groupList()[b]));
            g2g_ij.push_back(a * ng + b);
        }
    jvec("g2g", g2g);
    // i2i over a deterministic set of pairs, i2all and i2g for a few i
    std::vector<double> i2i, i2all, i2g;
    std::vector<int> pi, pj, ai, gi;
    int step = n > 40 ? n / 40 : 1;
    for (int i = 0; i < n; i += step)
        for (int j = 0; j < n; j += step + 1)
            if (i != j) { pi.push_back(i); pj.push_back(j); i2i.push_back(pot.i2i(spc.p, i, j)); }
    for (int i = 0; i < n; i += step) {
        ai.push_back(i);
        i2all.push_back(pot.i2all(spc.p, i));
        for (int g = 0; g < ng; ++g) { gi.push_back(g); i2g.push_back(pot.i2g(spc.p, *spc.groupList()[g], i)); }
    }
    jivec("i2i_i", pi); jivec("i2i_j", pj); jvec("i2i", i2i);
    jivec("i2all_i", ai); jvec("i2all", i2all);
    jivec("i2g_g", gi); jvec("i2g", i2g, false);
}

template<class Tenergy>
static int run_energy(Tmjson &in, const char *state)
{
    Tspace spc(in);
    Tenergy pot(in);
    spc.load(state, Tspace::RESIZE);
    std::printf("{\n");
    print_params(in);
    print_energies(spc, pot);
    std::printf("\n}\n");
    return 0;
}

template<class Tenergy>
static int run_mc(Tmjson &in, const char *state, int ncalls)
{
    Tspace spc(in);
    Tenergy pot(in);
    spc.load(state, Tspace::RESIZE);
    Recorder rec;
    // The moves section holds exactly one of "atomtranslate" / "moltransrot";
    // the order of construction and the engine seeding follow
    // Propagator::Propagator (move.h:7758-7858).
    std::unique_ptr<Move::Movebase<Tspace>> mv;
    auto m = in.at("moves");
    for (auto i = m.begin(); i != m.end(); ++i) {
        RecAT<Tspace>::internal().eng = slump.eng;   // move.h:7781
        if (i.key() == "atomtranslate") mv.reset(new RecAT<Tspace>(pot, spc, i.value(), &rec));
        if (i.key() == "moltransrot") mv.reset(new RecTR<Tspace>(pot, spc, i.value(), &rec));
    }
    if (!mv) { std::fprintf(stderr, "no move\n"); return 2; }
    double uinit = Energy::systemEnergy(spc, pot, spc.p);   // Propagator ufunction
    double utot = 0;
    for (int c = 0; c < ncalls; ++c) {
        // Propagator::move picks one of its moves with the internal engine
        (void)RecAT<Tspace>::internal().range(0, 0);
        utot += mv->move();
    }
    double ufinal = Energy::systemEnergy(spc, pot, spc.p);
    std::printf("{\n");
    print_params(in);
    jnum("uinit", uinit);
    jnum("ufinal", ufinal);
    jnum("utot", utot);
    jvec("du", rec.du);
    jivec("acc", rec.acc);
    jivec("who", rec.who);
    std::vector<double> xyz;
    for (auto &p : spc.p) { xyz.push_back(p.x()); xyz.push_back(p.y()); xyz.push_back(p.z()); }
    jvec("final_xyz", xyz);
    std::vector<double> cm;
    for (auto g : spc.groupList()) { cm.push_back(g->cm.x()); cm.push_back(g->cm.y()); cm.push_back(g->cm.z()); }
    jvec("final_cm", cm);
    std::ostringstream o;
    o << slump.eng;
    std::printf("\"final_rng\": \"%s\"\n}\n", o.str().c_str());
    return 0;
}

// Timed Metropolis loop for the CPU baseline: same construction as run_mc,
// but only the loop of `ncalls` move() calls is timed and no O(N^2) system
// energies are evaluated.  Prints trials, seconds and trials/s.
template<class Tenergy>
static int run_mctime(Tmjson &in, const char *state, int ncalls)
{
    Tspace spc(in);
    Tenergy pot(in);
    spc.load(state, Tspace::RESIZE);
    Recorder rec;
    std::unique_ptr<Move::Movebase<Tspace>> mv;
    auto m = in.at("moves");
    for (auto i = m.begin(); i != m.end(); ++i) {
        RecAT<Tspace>::internal().eng = slump.eng;
        if (i.key() == "atomtranslate") mv.reset(new RecAT<Tspace>(pot, spc, i.value(), &rec));
        if (i.key() == "moltransrot") mv.reset(new RecTR<Tspace>(pot, spc, i.value(), &rec));
    }
    if (!mv) return 2;
    auto t0 = std::chrono::steady_clock::now();
    double utot = 0;
    for (int c = 0; c < ncalls; ++c) {
        (void)RecAT<Tspace>::internal().range(0, 0);
        utot += mv->move();
    }
    auto t1 = std::chrono::steady_clock::now();
    double secs = std::chrono::duration<double>(t1 - t0).count();
    long acc = 0;
    for (int a : rec.acc) acc += a;
    std::printf("{\n");
    jnum("seconds", secs);
    jnum("trials", (double)rec.du.size());
    jnum("accepted", (double)acc);
    jnum("utot", utot);
    jnum("trials_per_s", rec.du.size() / secs, false);
    std::printf("\n}\n");
    return 0;
}

typedef CombinedPairPotential<DebyeHuckel, HardSphere> TDHHS;

// A pair Hamiltonian + ExternalPressure (energy.h:1817-1894) with a
// Propagator holding one Isobaric move (move.h:5424-5640): energies, or a
// recorded Metropolis run of `ncalls` calls.  With ncalls_at > 0 the
// Propagator also holds the input's AtomicTranslation move and the run goes on
// with `ncalls_at` calls of that move (so the structure factors an accepted
// volume move leaves behind reach the translation trials that follow).
template<class Tpair>
static int run_npt(const std::string &mode, Tmjson &in, const char *state, int ncalls,
                   int ncalls_at)
{
    Tspace spc(in);
    auto pot = Tpair(in) + Energy::ExternalPressure<Tspace>(in);
    spc.load(state, Tspace::RESIZE);
    const double P = in["moves"]["isobaric"].at("pressure").get<double>() * 1.0_mM;
    std::printf("{\n");
    print_params(in);
    jnum("pressure", P);
    if (mode == "energy") {
        print_energies(spc, pot);
        std::printf("\n}\n");
        return 0;
    }
    Recorder rec;
    // Propagator constructor (move.h:7758-7858): every move is built before
    // the first call (and each Movebase constructor calls setSpace,
    // move.h:403-406), the internal engine seeded from the global one
    std::unique_ptr<RecAT<Tspace>> at;
    if (ncalls_at > 0) {
        RecAT<Tspace>::internal().eng = slump.eng;
        at.reset(new RecAT<Tspace>(pot, spc, in["moves"]["atomtranslate"], &rec));
    }
    RecAT<Tspace>::internal().eng = slump.eng;   // move.h:7781
    RecIso<Tspace> mv(pot, spc, in["moves"]["isobaric"], &rec);
    double uinit = Energy::systemEnergy(spc, pot, spc.p);
    double utot = 0;
    for (int c = 0; c < ncalls; ++c) {
        (void)RecAT<Tspace>::internal().range(0, 0);
        utot += mv.move();
    }
    const size_t nvol = rec.du.size();
    double utot_vol = utot;
    for (int c = 0; c < ncalls_at; ++c) {
        (void)RecAT<Tspace>::internal().range(0, 0);
        utot += at->move();
    }
    double ufinal = Energy::systemEnergy(spc, pot, spc.p);
    jnum("uinit", uinit);
    jnum("ufinal", ufinal);
    jnum("utot", utot);
    jnum("utot_vol", utot_vol);
    jnum("nvol", (double)nvol);
    jvec("du", rec.du);
    jivec("acc", rec.acc);
    jivec("who", rec.who);
    std::vector<double> xyz;
    for (auto &p : spc.p) { xyz.push_back(p.x()); xyz.push_back(p.y()); xyz.push_back(p.z()); }
    jvec("final_xyz", xyz);
    std::vector<double> cm;
    for (auto g : spc.groupList()) { cm.push_back(g->cm.x()); cm.push_back(g->cm.y()); cm.push_back(g->cm.z()); }
    jvec("final_cm", cm);
    jvec("final_len", {spc.geo.len.x(), spc.geo.len.y(), spc.geo.len.z()});
    std::ostringstream o;
    o << slump.eng;
    std::printf("\"final_rng\": \"%s\"\n}\n", o.str().c_str());
    return 0;
}
// A pair Hamiltonian + ExternalPressure with one Propagator holding every
// move of the input that this path covers (atomtranslate, isobaric,
// moltransrot), built in the input's key order (nlohmann objects iterate
// sorted keys) and picked per call as Propagator::move picks them
// (move.h:7758-7858, 7877-7891: element() = begin + range(0, n - 1)).
template<class Tpair>
static int run_npt_prop(Tmjson &in, const char *state, int ncalls)
{
    Tspace spc(in);
    auto pot = Tpair(in) + Energy::ExternalPressure<Tspace>(in);
    spc.load(state, Tspace::RESIZE);
    Recorder rec;
    std::vector<std::unique_ptr<Move::Movebase<Tspace>>> mv;
    auto m = in.at("moves");
    for (auto i = m.begin(); i != m.end(); ++i) {
        RecAT<Tspace>::internal().eng = slump.eng;   // move.h:7781
        if (i.key() == "random" && i.value().is_object()) {   // move.h:7783-7787
            auto v = i.value();
            RecAT<Tspace>::internal() = RandomTwister<>(v);
        }
        if (i.key() == "atomtranslate") mv.emplace_back(new RecAT<Tspace>(pot, spc, i.value(), &rec));
        if (i.key() == "isobaric") mv.emplace_back(new RecIso<Tspace>(pot, spc, i.value(), &rec));
        if (i.key() == "moltransrot") mv.emplace_back(new RecTR<Tspace>(pot, spc, i.value(), &rec));
    }
    if (mv.empty()) return 2;
    std::vector<int> picks;
    double uinit = Energy::systemEnergy(spc, pot, spc.p);
    double utot = 0;
    for (int c = 0; c < ncalls; ++c) {
        int k = RecAT<Tspace>::internal().range(0, (int)mv.size() - 1);
        picks.push_back(k);
        utot += mv[k]->move();
    }
    double ufinal = Energy::systemEnergy(spc, pot, spc.p);
    std::printf("{\n");
    print_params(in);
    jnum("pressure", in["moves"]["isobaric"].at("pressure").get<double>() * 1.0_mM);
    jnum("uinit", uinit);
    jnum("ufinal", ufinal);
    jnum("utot", utot);
    jivec("picks", picks);
    jvec("du", rec.du);
    jivec("acc", rec.acc);
    jivec("who", rec.who);
    std::vector<double> xyz;
    for (auto &p : spc.p) { xyz.push_back(p.x()); xyz.push_back(p.y()); xyz.push_back(p.z()); }
    jvec("final_xyz", xyz);
    std::vector<double> cm;
    for (auto g : spc.groupList()) { cm.push_back(g->cm.x()); cm.push_back(g->cm.y()); cm.push_back(g->cm.z()); }
    jvec("final_cm", cm);
    jvec("final_len", {spc.geo.len.x(), spc.geo.len.y(), spc.geo.len.z()});
    std::ostringstream o;
    o << slump.eng;
    std::printf("\"final_rng\": \"%s\"\n}\n", o.str().c_str());
    return 0;
}

typedef CombinedPairPotential<Coulomb, HardSphere> TCoulHS;

int main(int argc, char **argv)
{
    std::cout.setstate(std::ios_base::failbit);   // silence library chatter on cout
    if (argc >= 4 && std::string(argv[1]) == "rng") {
        unsigned seed = (unsigned)std::strtoul(argv[2], nullptr, 10);
        int n = std::atoi(argv[3]);
        RandomTwister<> r;
        if (seed) r.seed(seed);
        std::printf("{\n\"u32\": [");
        for (int i = 0; i < n; ++i) std::printf("%lu%s", (unsigned long)r.eng(), i + 1 < n ? "," : "");
        std::printf("],\n");
        std::vector<double> can;
        for (int i = 0; i < n; ++i) can.push_back(r());
        jvec("canonical", can);
        std::vector<int> rng9, rng499, rng0;
        for (int i = 0; i < n; ++i) rng9.push_back(r.range(0, 9));
        for (int i = 0; i < n; ++i) rng499.push_back(r.range(0, 499));
        for (int i = 0; i < n; ++i) rng0.push_back(r.range(0, 0));
        jivec("range_0_9", rng9);
        jivec("range_0_499", rng499);
        jivec("range_0_0", rng0);
        std::vector<double> half;
        for (int i = 0; i < n; ++i) half.push_back(r.half());
        jvec("half", half, false);
        std::printf("\n}\n");
        return 0;
    }
    if (argc < 5) { std::fprintf(stderr, "usage: see header\n"); return 2; }
    std::string mode = argv[1], kind = argv[2];
    Tmjson in = openjson(argv[3]);
    const char *state = argv[4];
    int ncalls = argc > 5 ? std::atoi(argv[5]) : 1;
#define DISPATCH(E) return mode == "energy" ? run_energy<E>(in, state) : \
        mode == "mctime" ? run_mctime<E>(in, state, ncalls) : run_mc<E>(in, state, ncalls)
    typedef Energy::Nonbonded<Tspace, TDHHS> E_dhhs;
    typedef Energy::NonbondedVector<Tspace, TDHHS> E_dhhs_vec;
    typedef Energy::NonbondedCutg2g<Tspace, TDHHS> E_dhhs_cut;
    typedef Energy::Nonbonded<Tspace, TCoulHS> E_coulhs;
    typedef Energy::Nonbonded<Tspace, DebyeHuckel> E_dh;
    typedef Energy::Nonbonded<Tspace, LennardJonesLB> E_lj;
    typedef Energy::NonbondedEwald<Tspace, LennardJonesLB> E_ewaldlj;
    // cut and shifted LJ (energy.nonbonded.cutoff); NonbondedEwald cannot
    // take it (potentials.h:3770 hides the vector-distance overload)
    typedef Energy::Nonbonded<Tspace, CutShift<LennardJonesLB>> E_ljcut;
    if (kind == "dhhs") { DISPATCH(E_dhhs); }
    if (kind == "dhhs_vec") { DISPATCH(E_dhhs_vec); }
    if (kind == "dhhs_cut") { DISPATCH(E_dhhs_cut); }
    if (kind == "coulhs") { DISPATCH(E_coulhs); }
    if (kind == "dh") { DISPATCH(E_dh); }
    if (kind == "lj") { DISPATCH(E_lj); }
    if (kind == "ewaldlj") { DISPATCH(E_ewaldlj); }
    if (kind == "ljcut") { DISPATCH(E_ljcut); }
    const int ncalls_at = argc > 6 ? std::atoi(argv[6]) : 0;
    if (mode == "mcprop" && kind == "dhhs_npt") return run_npt_prop<E_dhhs>(in, state, ncalls);
    if (mode == "mcprop" && kind == "ewaldlj_npt") return run_npt_prop<E_ewaldlj>(in, state, ncalls);
    if (kind == "dhhs_npt") return run_npt<E_dhhs>(mode, in, state, ncalls, ncalls_at);
    if (kind == "ewaldlj_npt") return run_npt<E_ewaldlj>(mode, in, state, ncalls, ncalls_at);
    std::fprintf(stderr, "unknown kind %s\n", kind.c_str());
    return 2;
}
