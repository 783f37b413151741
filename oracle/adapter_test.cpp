// adapter_test.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Runs the reference's own Metropolis machinery three ways on the same
// input (reference JSON + Space::save state file) and prints the results:
//   A  reference CPU: Energy::Nonbonded / NonbondedEwald + AtomicTranslation
//      or TranslateRotate (the recorded trajectory)
//   B  the reference moves with include/faunus_amd/reference_adapter.h's
//      GPUNonbonded energy and GPU move subclasses (host-driven, one device
//      call per trial)
//   C  GPUPropagator: the whole Propagator::move on the device
// Built from the reference sources in place by oracle/Makefile (target
// adapter), output in oracle/_ref/, run by tests/test_gpu_adapter.py.
//
// Usage: adapter_test <kind> <config.json> <statefile> <ncalls> [ranks]
// (kinds ending in _npt: ExternalPressure and a Propagator holding every move
// of the input, Isobaric among them; dhhs_temper: `ranks` replicas of the
// input at Bjerrum lengths lB (1 + r/4), each with a GPUPropagator of the
// input's moves (a "temper" move among them), driven in lock-step by
// Move::GPUParallelTempering -- printed for the test to replay on the oracle)

#include <faunus/faunus.h>
#include <faunus/ewald.h>

#include <cstdio>
#include <memory>

#include "faunus_amd/reference_adapter.h"

using namespace Faunus;
using namespace Faunus::Potential;

typedef Space<Geometry::Cuboid, PointParticle> Tspace;

struct Rec {
    std::vector<double> du;
    std::vector<int> acc;
};

template <class Base>
struct Recording : public Base {
    Rec *rec;
    template <class E>
    Recording(E &e, Tspace &s, Tmjson &j, Rec *r) : Base(e, s, j), rec(r) {}
    static RandomTwister<> &internal() { return Move::Movebase<Tspace>::_slump(); }

protected:
    double _energyChange() override
    {
        double du = Base::_energyChange();
        rec->du.push_back(du);
        return du;
    }
    void _acceptMove() override { rec->acc.push_back(1); Base::_acceptMove(); }
    void _rejectMove() override { rec->acc.push_back(0); Base::_rejectMove(); }
};

static void pvec(const char *k, const std::vector<double> &v, bool comma = true)
{
    std::printf("\"%s\": [", k);
    for (size_t i = 0; i < v.size(); ++i) {
        if (std::isinf(v[i])) std::printf("\"%s\"", v[i] > 0 ? "inf" : "-inf");
        else if (std::isnan(v[i])) std::printf("\"nan\"");
        else std::printf("%.17g", v[i]);
        std::printf(i + 1 < v.size() ? "," : "");
    }
    std::printf("]%s\n", comma ? "," : "");
}

static void pivec(const char *k, const std::vector<int> &v)
{
    std::printf("\"%s\": [", k);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%d%s", v[i], i + 1 < v.size() ? "," : "");
    std::printf("],\n");
}

static std::vector<double> coords(Tspace &s)
{
    std::vector<double> x;
    for (auto &p : s.p) { x.push_back(p.x()); x.push_back(p.y()); x.push_back(p.z()); }
    return x;
}

static std::string engine_text()
{
    std::ostringstream o;
    o << slump.eng;
    return o.str();
}

template <class Tmove, class Tenergy>
static std::unique_ptr<Move::Movebase<Tspace>> make_move(Tmjson &in, Tenergy &pot, Tspace &spc, Rec *rec)
{
    std::unique_ptr<Move::Movebase<Tspace>> mv;
    auto m = in.at("moves");
    for (auto i = m.begin(); i != m.end(); ++i) {
        Recording<Move::AtomicTranslation<Tspace>>::internal().eng = slump.eng;
        mv.reset(new Tmove(pot, spc, i.value(), rec));
    }
    return mv;
}

template <class Eref>
static int run(const std::string &kind, Tmjson &in, const char *state, int ncalls, int elec, int sr,
               int nbkind, bool cutshift = false)
{
    bool tr = in.at("moves").count("moltransrot") > 0;
    std::printf("{\n");
    // ---- A: reference CPU
    {
        Tspace spc(in);
        Eref pot(in);
        spc.load(state, Tspace::RESIZE);
        Rec rec;
        std::unique_ptr<Move::Movebase<Tspace>> mv =
            tr ? make_move<Recording<Move::TranslateRotate<Tspace>>>(in, pot, spc, &rec)
               : make_move<Recording<Move::AtomicTranslation<Tspace>>>(in, pot, spc, &rec);
        for (int c = 0; c < ncalls; ++c) {
            (void)Recording<Move::AtomicTranslation<Tspace>>::internal().range(0, 0);
            mv->move();
        }
        pvec("A_du", rec.du);
        pivec("A_acc", rec.acc);
        pvec("A_xyz", coords(spc));
        std::printf("\"A_rng\": \"%s\",\n", engine_text().c_str());
    }
    // ---- B: reference moves, GPU energy
    {
        Tspace spc(in);
        spc.load(state, Tspace::RESIZE);
        auto prm = AMD::GPUParams::fromJson(in, elec, sr, nbkind, FGX_FP64, cutshift);
        Energy::GPUNonbonded<Tspace> pot(prm);
        Rec rec;
        std::unique_ptr<Move::Movebase<Tspace>> mv =
            tr ? make_move<Recording<Move::GPUTranslateRotate<Tspace>>>(in, pot, spc, &rec)
               : make_move<Recording<Move::GPUAtomicTranslation<Tspace>>>(in, pot, spc, &rec);
        for (int c = 0; c < ncalls; ++c) {
            (void)Recording<Move::AtomicTranslation<Tspace>>::internal().range(0, 0);
            mv->move();
        }
        pvec("B_du", rec.du);
        pivec("B_acc", rec.acc);
        pvec("B_xyz", coords(spc));
        std::printf("\"B_rng\": \"%s\",\n", engine_text().c_str());
    }
    // ---- C: device-resident Propagator
    {
        Tspace spc(in);
        spc.load(state, Tspace::RESIZE);
        auto prm = AMD::GPUParams::fromJson(in, elec, sr, nbkind, FGX_FP64, cutshift);
        Energy::GPUNonbonded<Tspace> pot(prm);
        Move::GPUPropagator<Tspace> prop(in, pot, spc);
        prop.move(ncalls);
        std::printf("\"C_trials\": %ld,\n\"C_accepted\": %ld,\n", prop.last.trials, prop.last.accepted);
        pvec("C_xyz", coords(spc));
        std::printf("\"C_rng\": \"%s\"\n}\n", engine_text().c_str());
    }
    return 0;
}

// The moves of the input in its key order (the Propagator's list,
// move.h:7789-7843), the engine seeded as the Propagator's constructor does.
template <class Tat, class Tiso, class Ttr, class Tenergy>
static std::vector<std::unique_ptr<Move::Movebase<Tspace>>> make_moves(Tmjson &in, Tenergy &pot,
                                                                       Tspace &spc, Rec *rec)
{
    std::vector<std::unique_ptr<Move::Movebase<Tspace>>> mv;
    auto m = in.at("moves");
    for (auto i = m.begin(); i != m.end(); ++i) {
        Recording<Move::AtomicTranslation<Tspace>>::internal().eng = slump.eng;
        if (i.key() == "random" && i.value().is_object()) {   // move.h:7783-7787
            auto v = i.value();
            Recording<Move::AtomicTranslation<Tspace>>::internal() = RandomTwister<>(v);
        }
        if (i.key() == "atomtranslate") mv.emplace_back(new Tat(pot, spc, i.value(), rec));
        if (i.key() == "isobaric") mv.emplace_back(new Tiso(pot, spc, i.value(), rec));
        if (i.key() == "moltransrot") mv.emplace_back(new Ttr(pot, spc, i.value(), rec));
    }
    return mv;
}

static std::vector<double> box_of(Tspace &s)
{
    return {s.geo.len.x(), s.geo.len.y(), s.geo.len.z()};
}

// A Hamiltonian with Energy::ExternalPressure and a Propagator holding every
// move of the input (Isobaric among them), run the three ways
template <class Eref>
static int run_npt(Tmjson &in, const char *state, int ncalls, int elec, int sr, int nbkind)
{
    std::printf("{\n");
    // ---- A: reference CPU
    {
        Tspace spc(in);
        auto pot = Eref(in) + Energy::ExternalPressure<Tspace>(in);
        spc.load(state, Tspace::RESIZE);
        Rec rec;
        auto mv = make_moves<Recording<Move::AtomicTranslation<Tspace>>,
                             Recording<Move::Isobaric<Tspace>>,
                             Recording<Move::TranslateRotate<Tspace>>>(in, pot, spc, &rec);
        (void)Energy::systemEnergy(spc, pot, spc.p);   // Propagator's uinit
        for (int c = 0; c < ncalls; ++c) {
            int k = Recording<Move::AtomicTranslation<Tspace>>::internal().range(0, (int)mv.size() - 1);
            mv[k]->move();
        }
        pvec("A_du", rec.du);
        pivec("A_acc", rec.acc);
        pvec("A_xyz", coords(spc));
        pvec("A_len", box_of(spc));
        std::printf("\"A_rng\": \"%s\",\n", engine_text().c_str());
    }
    // ---- B: reference moves, GPU energy (the reference's ExternalPressure
    // term beside it in the Hamiltonian)
    {
        Tspace spc(in);
        spc.load(state, Tspace::RESIZE);
        auto prm = AMD::GPUParams::fromJson(in, elec, sr, nbkind, FGX_FP64);
        auto &pot = Energy::GPUNonbonded<Tspace>(prm) + Energy::ExternalPressure<Tspace>(in);
        Rec rec;
        auto mv = make_moves<Recording<Move::GPUAtomicTranslation<Tspace>>,
                             Recording<Move::GPUIsobaric<Tspace>>,
                             Recording<Move::GPUTranslateRotate<Tspace>>>(in, pot, spc, &rec);
        (void)Energy::systemEnergy(spc, pot, spc.p);
        for (int c = 0; c < ncalls; ++c) {
            int k = Recording<Move::AtomicTranslation<Tspace>>::internal().range(0, (int)mv.size() - 1);
            mv[k]->move();
        }
        pvec("B_du", rec.du);
        pivec("B_acc", rec.acc);
        pvec("B_xyz", coords(spc));
        pvec("B_len", box_of(spc));
        std::printf("\"B_rng\": \"%s\",\n", engine_text().c_str());
    }
    // ---- C: device-resident Propagator (ExternalPressure on the device)
    {
        Tspace spc(in);
        spc.load(state, Tspace::RESIZE);
        auto prm = AMD::GPUParams::fromJson(in, elec, sr, nbkind, FGX_FP64, false, true);
        Energy::GPUNonbonded<Tspace> pot(prm);
        Move::GPUPropagator<Tspace> prop(in, pot, spc);
        prop.move(ncalls);
        std::printf("\"C_trials\": %ld,\n\"C_accepted\": %ld,\n", prop.last.trials,
                    prop.last.accepted);
        pvec("C_xyz", coords(spc));
        pvec("C_len", box_of(spc));
        std::printf("\"C_rng\": \"%s\"\n}\n", engine_text().c_str());
    }
    return 0;
}

// D: replicas on one GPU, Move::GPUParallelTempering over their GPUPropagators
static int run_temper(Tmjson &in, const char *state, int ncalls, int R)
{
    std::string eng0;   // the state file's engine: every rank starts from it
    std::vector<std::unique_ptr<Tspace>> spc;
    std::vector<std::unique_ptr<Energy::GPUNonbonded<Tspace>>> pot;
    std::vector<std::unique_ptr<Move::GPUPropagator<Tspace>>> prop;
    std::vector<Move::GPUPropagator<Tspace> *> ranks;
    auto prm0 = AMD::GPUParams::fromJson(in, FGX_ELEC_DH, FGX_SR_HS, FGX_NB_PLAIN, FGX_FP64);
    // every Space first: each one reads the input's atom list into the
    // global atom table, which the device systems then copy
    for (int r = 0; r < R; ++r) {
        spc.emplace_back(new Tspace(in));
        spc[r]->load(state, Tspace::RESIZE);
        if (r == 0) eng0 = engine_text();
    }
    for (int r = 0; r < R; ++r) {
        auto prm = prm0;
        prm.lB = prm0.lB * (1.0 + 0.25 * r);
        pot.emplace_back(new Energy::GPUNonbonded<Tspace>(prm));
        std::istringstream is(eng0);
        is >> slump.eng;
        prop.emplace_back(new Move::GPUPropagator<Tspace>(in, *pot[r], *spc[r]));
        ranks.push_back(prop[r].get());
    }
    Move::GPUParallelTempering<Tspace> pt(ranks, 11u);
    pt.propagate(ncalls);
    std::printf("{\n\"D_lB0\": %.17g,\n\"D_rng0\": \"%s\",\n", prm0.lB, eng0.c_str());
    std::vector<double> calls, du, ret;
    std::vector<int> acc;
    for (auto &h : pt.history) {
        calls.push_back((double)h.call);
        for (int r = 0; r < R; ++r) {
            acc.push_back(h.acc[r]);
            du.push_back(h.du[r]);
            ret.push_back(h.ret[r]);
        }
    }
    pvec("D_calls", calls);
    pivec("D_acc", acc);
    pvec("D_du", du);
    pvec("D_ret", ret);
    std::vector<double> x;
    for (int r = 0; r < R; ++r) {
        auto c = coords(*spc[r]);
        x.insert(x.end(), c.begin(), c.end());
    }
    pvec("D_xyz", x);
    std::printf("\"D_swaps\": %d\n}\n", (int)pt.accmap.size());
    return 0;
}

int main(int argc, char **argv)
{
    std::cout.setstate(std::ios_base::failbit);
    if (argc < 5) { std::fprintf(stderr, "usage: adapter_test kind config state ncalls\n"); return 2; }
    std::string kind = argv[1];
    Tmjson in = openjson(argv[2]);
    int ncalls = std::atoi(argv[4]);
    typedef Energy::Nonbonded<Tspace, CombinedPairPotential<DebyeHuckel, HardSphere>> E_dhhs;
    typedef Energy::NonbondedCutg2g<Tspace, CombinedPairPotential<DebyeHuckel, HardSphere>> E_dhhs_cut;
    typedef Energy::Nonbonded<Tspace, CombinedPairPotential<Coulomb, HardSphere>> E_coulhs;
    typedef Energy::Nonbonded<Tspace, LennardJonesLB> E_lj;
    typedef Energy::NonbondedEwald<Tspace, LennardJonesLB> E_ewaldlj;
    typedef Energy::Nonbonded<Tspace, CutShift<LennardJonesLB>> E_ljcut;
    if (kind == "dhhs") return run<E_dhhs>(kind, in, argv[3], ncalls, FGX_ELEC_DH, FGX_SR_HS, FGX_NB_PLAIN);
    if (kind == "dhhs_cut") return run<E_dhhs_cut>(kind, in, argv[3], ncalls, FGX_ELEC_DH, FGX_SR_HS, FGX_NB_CUTG2G);
    if (kind == "coulhs") return run<E_coulhs>(kind, in, argv[3], ncalls, FGX_ELEC_COULOMB, FGX_SR_HS, FGX_NB_PLAIN);
    if (kind == "lj") return run<E_lj>(kind, in, argv[3], ncalls, FGX_ELEC_NONE, FGX_SR_LJ, FGX_NB_PLAIN);
    if (kind == "ewaldlj") return run<E_ewaldlj>(kind, in, argv[3], ncalls, FGX_ELEC_EWALD, FGX_SR_LJ, FGX_NB_VECTOR);
    if (kind == "ljcut") return run<E_ljcut>(kind, in, argv[3], ncalls, FGX_ELEC_NONE, FGX_SR_LJ, FGX_NB_PLAIN, true);
    if (kind == "dhhs_temper")
        return run_temper(in, argv[3], ncalls, argc > 5 ? std::atoi(argv[5]) : 4);
    if (kind == "dhhs_npt")
        return run_npt<E_dhhs>(in, argv[3], ncalls, FGX_ELEC_DH, FGX_SR_HS, FGX_NB_PLAIN);
    if (kind == "ewaldlj_npt")
        return run_npt<E_ewaldlj>(in, argv[3], ncalls, FGX_ELEC_EWALD, FGX_SR_LJ, FGX_NB_VECTOR);
    std::fprintf(stderr, "unknown kind\n");
    return 2;
}
