// sg_ga.cpp -- the genetic algorithm that drives a FitSKIRT fit.
//
// FitSKIRT's Optimization (FitSKIRTcore/Optimization.cpp:42-90) runs GAlib's steady-state genetic
// algorithm over one real gene per parameter range: uniform initialisation, Gaussian mutation, uniform
// crossover, minimisation with sigma-truncation scaling and roulette-wheel selection; OligoFitScheme::runSelf
// (OligoFitScheme.cpp:27-34) calls initialize() and then step() until done(). Each evaluation of the
// population hands the individuals that are not yet evaluated to PopEvaluate (Optimization.cpp:301-333),
// which runs one full simulation per individual. Here the same individuals are handed to the caller as one
// batch (sg_ga_pending), who evaluates them together on the GPU and returns their scores
// (sg_ga_set_scores); initialize and step each end at that hand-over and finish when the scores arrive.
//
// The algorithm makes GAlib's choices draw for draw, so that a fit follows the reference's trajectory for
// the same seed and the same scores:
//  - uniform deviates from Numerical Recipes' ran2 (GAlib/garandom.cpp:255-330), seeded as GARandomSeed
//    (garandom.cpp:61-83; seed 0 takes one from the clock and process id); random bits from NR's
//    primitive-polynomial shift register (garandom.cpp:160-185) seeded with the same value; unit Gaussians
//    by the polar Box-Muller method with the second deviate kept for the next call (garandom.cpp:118-140);
//    integer draws lo + (int)((hi - lo + 1) u), coin flips u <= p (p = 0 and 1 draw nothing) (garandom.h);
//  - the population keeps two orders of the same individuals, by raw score and by scaled fitness, and
//    refreshes statistics, scaling, sorting and the selection table only when asked for and out of date
//    (GAPopulation.cpp); sums run in those orders, and the sorts are GAlib's quicksort (pivot on the right,
//    GAPopulation.cpp:560-620), whose order among equal keys the roulette wheel and the removal of the
//    worst individuals depend on;
//  - setting a gene to a different value marks the individual unevaluated; copying an individual copies
//    its score and evaluated flag (GA1DArrayGenome.h:82-84, GAGenome.cpp:55-68), so only children that
//    crossover or mutation changed are evaluated again.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <ctime>
#include <memory>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/skirt7_amd.h"

namespace sg {
void set_error(const std::string& msg);
}

namespace {

// ---- random numbers ----------------------------------------------------------------------------------
class GaRandom
{
public:
    void seed(uint32_t s)
    {
        // ran2's seeding: 8 warm-up steps of the first generator, then the shuffle table
        int64_t id = (int64_t)s;
        if (id == 0) id = 1;
        if (id < 0) id = -id;
        id2_ = id;
        for (int j = kTab + 7; j >= 0; j--) {
            id = lcg(id, kA1, kQ1, kR1, kM1);
            if (j < kTab) tab_[j] = id;
        }
        id1_ = id;
        out_ = tab_[0];
        bits_ = s;
        cached_ = false;
    }

    double uniform()
    {
        id1_ = lcg(id1_, kA1, kQ1, kR1, kM1);
        id2_ = lcg(id2_, kA2, kQ2, kR2, kM2);
        int j = (int)(out_ / (1 + (kM1 - 1) / kTab));
        out_ = tab_[j] - id2_;
        tab_[j] = id1_;
        if (out_ < 1) out_ += kM1 - 1;
        double u = (1.0 / kM1) * out_;
        const double umax = 1.0 - 1.2e-7;
        return u > umax ? umax : u;
    }

    double uniform(double lo, double hi)
    {
        double v = hi - lo;
        v *= uniform();
        return v + lo;
    }

    int integer(int lo, int hi)
    {
        double v = (double)hi - (double)lo + 1.0;
        v *= uniform();
        return (int)v + lo;
    }

    bool coin(double p)
    {
        if (p == 1.0) return true;
        if (p == 0.0) return false;
        return uniform() <= p;
    }

    int bit()
    {
        // x^18 + x^5 + x^2 + x + 1 shift register; only bit 18 is tested, so the word may overflow freely
        const uint64_t top = 1u << 17, taps = 1u | 2u | 16u;
        if (bits_ & top) {
            bits_ = ((bits_ ^ taps) << 1) | 1u;
            return 1;
        }
        bits_ <<= 1;
        return 0;
    }

    double gaussian()
    {
        if (cached_) {
            cached_ = false;
            return cache_;
        }
        double a, b, r2;
        do {
            a = 2.0 * uniform() - 1.0;
            b = 2.0 * uniform() - 1.0;
            r2 = a * a + b * b;
        } while (r2 >= 1.0 || r2 == 0.0);
        double v = -2.0 * std::log(r2) / r2;
        double f = v > 0.0 ? std::sqrt(v) : 0.0;
        cache_ = a * f;
        cached_ = true;
        return b * f;
    }

private:
    static const int64_t kM1 = 2147483563, kM2 = 2147483399, kA1 = 40014, kA2 = 40692;
    static const int64_t kQ1 = 53668, kQ2 = 52774, kR1 = 12211, kR2 = 3791;
    static const int kTab = 32;

    // Schrage's multiplication a*x mod m without overflow
    static int64_t lcg(int64_t x, int64_t a, int64_t q, int64_t r, int64_t m)
    {
        int64_t k = x / q;
        x = a * (x - k * q) - k * r;
        return x < 0 ? x + m : x;
    }

    int64_t id1_ = 0, id2_ = 123456789, out_ = 0, tab_[kTab] = {};
    uint64_t bits_ = 0;
    bool cached_ = false;
    double cache_ = 0.0;
};

// ---- individuals and population --------------------------------------------------------------------
struct Individual
{
    std::vector<double> gene;
    double score = 0.0, fitness = 0.0;
    bool evaluated = false;

    void set(int k, double v)
    {
        if (gene[k] != v) {
            gene[k] = v;
            evaluated = false;
        }
    }
    void copy_from(const Individual& o)
    {
        if (&o == this) return;
        gene = o.gene;
        score = o.score;
        fitness = o.fitness;
        evaluated = o.evaluated;
    }
};

enum class Phase { Created, Initializing, Idle, Stepping };

// Quicksort ascending on key(), partitioning around the rightmost element exactly as GAlib's population
// sorts do (the scan from the right stops at index 0, not at the subarray's left end: the element left of
// a right-hand subarray is never larger than its keys, so it stops the scan too).
template <class Key>
void quicksort(std::vector<Individual*>& c, int l, int r, Key key)
{
    if (r <= l) return;
    double v = key(c[r]);
    int i = l - 1, j = r;
    for (;;) {
        while (key(c[++i]) < v && i <= r) {
        }
        while (key(c[--j]) > v && j > 0) {
        }
        if (i >= j) break;
        std::swap(c[i], c[j]);
    }
    std::swap(c[i], c[r]);
    quicksort(c, l, i - 1, key);
    quicksort(c, i + 1, r, key);
}

struct Ga
{
    int L = 0, popsize = 0, ngen = 0, nrepl = 0;
    double pmut = 0.0, pcross = 0.0;
    uint32_t seed = 0;
    std::vector<double> lo, hi;
    GaRandom rng;

    std::vector<std::unique_ptr<Individual>> store;   // owns every individual (population and children)
    std::vector<Individual*> raw, scaled;              // the population in raw-score and in scaled order
    std::vector<Individual*> children;                 // the steady-state GA's replacement population
    std::vector<Individual*> pending;                  // individuals waiting for their scores

    bool rsorted = false, ssorted = false, is_scaled = false, statted = false, selectready = false;
    double rawsum = 0, rawave = 0, rawdev = 0, rawmax = 0, rawmin = 0;
    double fitsum = 0, fitave = 0, fitdev = 0, fitmax = 0, fitmin = 0;
    std::vector<double> psum;

    Phase phase = Phase::Created;
    int generation = 0;
    Individual best;          // GAlib's best-of-all record (one genome)
    double best_eval = 1e20;  // Optimization::_bestChi2 (Optimization.cpp:37,326-331)
    std::vector<double> best_eval_gene;

    Individual* fresh()
    {
        store.emplace_back(new Individual);
        store.back()->gene.assign(L, 0.0);
        return store.back().get();
    }

    int size() const { return (int)raw.size(); }

    void touch() { rsorted = ssorted = is_scaled = statted = selectready = false; }

    // GAPopulation::statistics (GAPopulation.cpp:236-268): raw-score moments in raw order
    void statistics()
    {
        if (statted) return;
        int n = size();
        double sum = raw[0]->score;
        rawmin = rawmax = sum;
        for (int i = 1; i < n; i++) {
            double s = raw[i]->score;
            sum += s;
            rawmax = rawmax > s ? rawmax : s;
            rawmin = rawmin < s ? rawmin : s;
        }
        rawave = sum / n;
        rawsum = sum;
        double var = 0.0;
        if (n > 1) {
            for (int i = 0; i < n; i++) {
                double d = raw[i]->score - rawave;
                var += d * d;
            }
            var /= (n - 1);
        }
        rawdev = std::sqrt(var);
        statted = true;
    }

    // GASigmaTruncationScaling::evaluate (GAScaling.cpp:116-124, multiplier 2) and the fitness moments of
    // GAPopulation::scale (GAPopulation.cpp:269-300) in scaled order
    void scale()
    {
        if (is_scaled) return;
        int n = size();
        for (int i = 0; i < n; i++) {
            statistics();
            double f = raw[i]->score - rawave;
            f += 2.0 * rawdev;
            if (f < 0) f = 0.0;
            raw[i]->fitness = f;
        }
        double sum = scaled[0]->fitness;
        fitmin = fitmax = sum;
        for (int i = 1; i < n; i++) {
            double f = scaled[i]->fitness;
            sum += f;
            fitmax = fitmax > f ? fitmax : f;
            fitmin = fitmin < f ? fitmin : f;
        }
        fitave = sum / n;
        fitsum = sum;
        double var = 0.0;
        if (n > 1) {
            for (int i = 0; i < n; i++) {
                double d = scaled[i]->fitness - fitave;
                var += d * d;
            }
            var /= (n - 1);
        }
        fitdev = std::sqrt(var);
        is_scaled = true;
        ssorted = false;
    }

    // GAPopulation::sort (GAPopulation.cpp:214-235), lowest first (the fit minimises)
    void sort_raw()
    {
        if (!rsorted) {
            quicksort(raw, 0, size() - 1, [](const Individual* g) { return g->score; });
            selectready = false;
        }
        rsorted = true;
    }
    void sort_scaled()
    {
        if (!ssorted) {
            quicksort(scaled, 0, size() - 1, [](const Individual* g) { return g->fitness; });
            selectready = false;
        }
        ssorted = true;
    }

    // GARouletteWheelSelector::update on scaled fitness (GASelector.cpp:108-175), lowest fitness best
    void prepselect()
    {
        if (selectready) return;
        int n = size();
        psum.assign(n, 0.0);
        scale();
        if (fitmax == fitmin) {
            for (int i = 0; i < n; i++) psum[i] = (double)(i + 1) / (double)n;
        } else if ((fitmax > 0 && fitmin >= 0) || (fitmax <= 0 && fitmin < 0)) {
            sort_scaled();
            psum[0] = -scaled[0]->fitness + fitmax + fitmin;
            for (int i = 1; i < n; i++) psum[i] = -scaled[i]->fitness + fitmax + fitmin + psum[i - 1];
            for (int i = 0; i < n; i++) psum[i] /= psum[n - 1];
        }
        selectready = true;
    }

    // GARouletteWheelSelector::select (GASelector.cpp:79-100)
    Individual* select()
    {
        prepselect();
        double cutoff = rng.uniform();
        int lower = 0, upper = size() - 1;
        while (upper >= lower) {
            int i = lower + (upper - lower) / 2;
            if (psum[i] > cutoff)
                upper = i - 1;
            else
                lower = i + 1;
        }
        lower = std::min(size() - 1, lower);
        lower = std::max(0, lower);
        return scaled[lower];
    }

    // GARealGaussianMutator (GARealGenome.h:165-205): bounded genes move by a Gaussian of 5% of the range
    void mutate(Individual* c)
    {
        if (pmut <= 0.0) return;
        double nmut = pmut * (double)L;
        if (nmut < 1.0) {
            for (int i = L - 1; i >= 0; i--) {
                double v = c->gene[i];
                if (rng.coin(pmut)) {
                    v += rng.gaussian() * (0.05 * (hi[i] - lo[i]));
                    v = lo[i] > v ? lo[i] : v;
                    v = hi[i] < v ? hi[i] : v;
                    c->set(i, v);
                }
            }
        } else {
            for (int n = 0; n < nmut; n++) {
                int k = rng.integer(0, L - 1);
                double v = c->gene[k];
                v += rng.gaussian() * (0.05 * (hi[k] - lo[k]));
                v = lo[k] > v ? lo[k] : v;
                v = hi[k] < v ? hi[k] : v;
                c->set(k, v);
            }
        }
    }

    // the population evaluator: every individual that is not evaluated, in raw order
    void collect_pending()
    {
        pending.clear();
        for (Individual* g : raw)
            if (!g->evaluated) pending.push_back(g);
    }

    // GASteadyStateGA::initialize (GASStateGA.cpp:258-268) up to the evaluation
    void initialize()
    {
        uint32_t s = seed;
        if (s == 0) {
            unsigned long t = (unsigned long)std::time(nullptr) * (unsigned long)getpid();
            for (unsigned i = 0; i < 32 && s == 0; i++) s += (uint32_t)(t & (1ul << i));
            if (s == 0) s = 1;
        }
        rng.seed(s);
        for (Individual* g : raw) {
            // GA1DArrayAlleleGenome::UniformInitializer: resize to a random length in [L, L] (one draw),
            // then one uniform draw per gene, last gene first
            g->evaluated = false;
            (void)rng.integer(L, L);
            for (int k = L - 1; k >= 0; k--) g->set(k, rng.uniform(lo[k], hi[k]));
        }
        touch();
        collect_pending();
        phase = Phase::Initializing;
    }

    // GAStatistics::reset (GAStatistics.cpp:254-290), for what it changes in the population
    void finish_initialize()
    {
        touch();
        statistics();
        sort_raw();
        best.copy_from(*raw[0]);
        generation = 0;
        phase = Phase::Idle;
    }

    // GASteadyStateGA::step (GASStateGA.cpp:277-343) up to the evaluation
    void step()
    {
        int m = (int)children.size(), i;
        for (i = 0; i < m - 1; i += 2) {
            Individual* mom = select();
            Individual* dad = select();
            Individual* sis = children[i];
            Individual* bro = children[i + 1];
            if (rng.coin(pcross)) {
                // GA1DArrayGenome::UniformCrossover (GA1DArrayGenome.h:612-640), last gene first
                for (int k = L - 1; k >= 0; k--) {
                    if (rng.bit()) {
                        sis->set(k, mom->gene[k]);
                        bro->set(k, dad->gene[k]);
                    } else {
                        sis->set(k, dad->gene[k]);
                        bro->set(k, mom->gene[k]);
                    }
                }
            } else {
                sis->copy_from(*mom);
                bro->copy_from(*dad);
            }
            mutate(sis);
            mutate(bro);
        }
        if (m % 2 != 0) {
            Individual* mom = select();
            Individual* dad = select();
            Individual* sis = children[i];
            if (rng.coin(pcross)) {
                for (int k = L - 1; k >= 0; k--) sis->set(k, rng.bit() ? mom->gene[k] : dad->gene[k]);
            } else {
                sis->copy_from(rng.bit() ? *mom : *dad);
            }
            mutate(sis);
        }
        // GAPopulation::add: appended to both orders
        for (Individual* c : children) {
            raw.push_back(c);
            scaled.push_back(c);
        }
        touch();
        collect_pending();
        phase = Phase::Stepping;
    }

    // the rest of GASteadyStateGA::step: scale, remove the worst (scaled order) individuals into the
    // replacement population, then GAStatistics::update (GAStatistics.cpp:230-250)
    void finish_step()
    {
        touch();
        scale();
        for (size_t r = 0; r < children.size(); r++) {
            sort_scaled();
            int n = size();
            Individual* worst = scaled[n - 1];
            scaled.pop_back();
            raw = scaled;
            rsorted = false;
            is_scaled = statted = selectready = false;
            children[r] = worst;
        }
        generation++;
        statistics();
        sort_raw();
        if (raw[0]->score < best.score) best.copy_from(*raw[0]);
        phase = Phase::Idle;
    }
};

Ga* checked(sg_ga* h)
{
    if (!h) sg::set_error("sg_ga: null handle");
    return reinterpret_cast<Ga*>(h);
}

} // namespace

extern "C" {

int sg_ga_create(int ngenes, const double* lower, const double* upper, int popsize, int generations, double pmut,
                 double pcross, uint32_t seed, sg_ga** out)
{
    if (!out || ngenes < 1 || !lower || !upper || popsize < 1 || generations < 0 || !(pmut >= 0.0 && pmut <= 1.0) ||
        !(pcross >= 0.0 && pcross <= 1.0)) {
        sg::set_error("sg_ga_create: need at least one gene, bounds, a positive population size, a non-negative "
                      "generation count and probabilities in [0, 1]");
        return -1;
    }
    for (int k = 0; k < ngenes; k++) {
        if (!(lower[k] <= upper[k])) {
            sg::set_error("sg_ga_create: parameter range " + std::to_string(k) + " has its minimum above its maximum");
            return -1;
        }
    }
    std::unique_ptr<Ga> ga(new Ga);
    ga->L = ngenes;
    ga->lo.assign(lower, lower + ngenes);
    ga->hi.assign(upper, upper + ngenes);
    ga->popsize = popsize;
    ga->ngen = generations;
    ga->pmut = pmut;
    ga->pcross = pcross;
    ga->seed = seed;
    // GASteadyStateGA::populationSize with the default replacement fraction 0.5 (GABaseGA.cpp:24)
    double n = 0.5 * (double)popsize;
    ga->nrepl = n < 1 ? 1 : (int)n;
    for (int i = 0; i < popsize; i++) ga->raw.push_back(ga->fresh());
    ga->scaled = ga->raw;
    for (int i = 0; i < ga->nrepl; i++) ga->children.push_back(ga->fresh());
    ga->best.gene.assign(ngenes, 0.0);
    *out = reinterpret_cast<sg_ga*>(ga.release());
    return 0;
}

int sg_ga_destroy(sg_ga* h)
{
    delete reinterpret_cast<Ga*>(h);
    return 0;
}

int sg_ga_initialize(sg_ga* h)
{
    Ga* ga = checked(h);
    if (!ga) return -1;
    if (ga->phase == Phase::Initializing || ga->phase == Phase::Stepping) {
        sg::set_error("sg_ga_initialize: the previous evaluation has not received its scores");
        return -1;
    }
    ga->initialize();
    return 0;
}

int sg_ga_step(sg_ga* h)
{
    Ga* ga = checked(h);
    if (!ga) return -1;
    if (ga->phase != Phase::Idle) {
        sg::set_error(ga->phase == Phase::Created ? "sg_ga_step: call sg_ga_initialize first"
                                                  : "sg_ga_step: the previous evaluation has not received its scores");
        return -1;
    }
    ga->step();
    return 0;
}

int sg_ga_pending(sg_ga* h, int* n, double* genes)
{
    Ga* ga = checked(h);
    if (!ga) return -1;
    if (!n) {
        sg::set_error("sg_ga_pending: null count");
        return -1;
    }
    *n = (int)ga->pending.size();
    if (genes)
        for (size_t i = 0; i < ga->pending.size(); i++)
            std::memcpy(genes + i * ga->L, ga->pending[i]->gene.data(), sizeof(double) * ga->L);
    return 0;
}

int sg_ga_set_scores(sg_ga* h, int n, const double* scores)
{
    Ga* ga = checked(h);
    if (!ga) return -1;
    if (ga->phase != Phase::Initializing && ga->phase != Phase::Stepping) {
        sg::set_error("sg_ga_set_scores: no evaluation is pending");
        return -1;
    }
    if (n != (int)ga->pending.size() || (n > 0 && !scores)) {
        sg::set_error("sg_ga_set_scores: expected " + std::to_string(ga->pending.size()) + " scores, got " +
                      std::to_string(n));
        return -1;
    }
    for (int i = 0; i < n; i++) {
        ga->pending[i]->score = scores[i];
        ga->pending[i]->evaluated = true;
        if (scores[i] < ga->best_eval) {
            ga->best_eval = scores[i];
            ga->best_eval_gene = ga->pending[i]->gene;
        }
    }
    ga->pending.clear();
    if (ga->phase == Phase::Initializing)
        ga->finish_initialize();
    else
        ga->finish_step();
    return 0;
}

int sg_ga_state(sg_ga* h, int* generation, int* done)
{
    Ga* ga = checked(h);
    if (!ga) return -1;
    if (generation) *generation = ga->generation;
    if (done) *done = (ga->phase == Phase::Idle && ga->generation >= ga->ngen) ? 1 : 0;
    return 0;
}

int sg_ga_best(sg_ga* h, double* genes, double* score)
{
    Ga* ga = checked(h);
    if (!ga) return -1;
    if (ga->phase == Phase::Created || ga->phase == Phase::Initializing) {
        sg::set_error("sg_ga_best: the population has not been evaluated");
        return -1;
    }
    if (genes) std::memcpy(genes, ga->best.gene.data(), sizeof(double) * ga->L);
    if (score) *score = ga->best.score;
    return 0;
}

int sg_ga_population(sg_ga* h, double* genes, double* scores)
{
    Ga* ga = checked(h);
    if (!ga) return -1;
    for (int i = 0; i < ga->size(); i++) {
        if (genes) std::memcpy(genes + (size_t)i * ga->L, ga->raw[i]->gene.data(), sizeof(double) * ga->L);
        if (scores) scores[i] = ga->raw[i]->score;
    }
    return 0;
}

} // extern "C"
