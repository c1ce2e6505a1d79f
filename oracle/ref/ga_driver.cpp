// Test infrastructure: the checker for the product's genetic algorithm (skirt7_amd/csrc/sg_ga.cpp).
//
// Runs GAlib, compiled from the reference's own GAlib sources where they lie, set up exactly as FitSKIRT's
// Optimization::setupSelfBefore does (FitSKIRTcore/Optimization.cpp:42-90): a real genome with one bounded
// allele set per parameter range, uniform initializer, Gaussian mutator, uniform crossover, a steady-state GA
// that minimises with sigma-truncation scaling, score and flush frequencies 0; then OligoFitScheme::runSelf
// (OligoFitScheme.cpp:27-34): initialize(seed), step() until done(). The population evaluator plays the part
// of Optimization::PopEvaluate (Optimization.cpp:301-333): it collects the individuals that are not evaluated,
// in population order, and asks for their scores over a pipe.
//
// Protocol (stdin/stdout, one item per line, every double as a C99 hexadecimal float):
//   in : "L popsize generations seed", "pmut pcross", then L lines "lower upper"
//   out: "EVAL n generation" then n lines of L genes; in: n lines with one score each
//   out: "DONE generation", popsize lines "genes... score" (population order), "BEST genes... score"
// One GA run per process: GAlib keeps its random state (and the cached second Gaussian deviate) in globals.
#include <GARealGenome.h>
#include <GASStateGA.h>
#include <GAScaling.h>
#include <garandom.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

static int g_len = 0;

static double read_double()
{
    char buf[128];
    if (std::scanf("%127s", buf) != 1) std::exit(3);
    return std::strtod(buf, nullptr);
}

static void print_genes(const GAGenome& g)
{
    const GARealGenome& r = (const GARealGenome&)g;
    for (int k = 0; k < g_len; k++) std::printf("%s%a", k ? " " : "", r.gene(k));
}

static void pipe_evaluator(GAPopulation& p)
{
    std::vector<int> todo;
    for (int i = 0; i < p.size(); i++)
        if (p.individual(i).isEvaluated() == gaFalse) todo.push_back(i);
    int gen = p.geneticAlgorithm() ? p.geneticAlgorithm()->statistics().generation() : 0;
    std::printf("EVAL %d %d\n", (int)todo.size(), gen);
    for (int i : todo)
    {
        print_genes(p.individual(i));
        std::printf("\n");
    }
    std::fflush(stdout);
    for (int i : todo) p.individual(i).score(read_double());
}

int main()
{
    int popsize, generations;
    unsigned int seed;
    if (std::scanf("%d %d %d %u", &g_len, &popsize, &generations, &seed) != 4) return 2;
    double pmut = read_double();
    double pcross = read_double();
    GARealAlleleSetArray alleles;
    for (int k = 0; k < g_len; k++)
    {
        double lo = read_double();
        double hi = read_double();
        alleles.add(GARealAlleleSet(lo, hi));
    }

    GARealGenome genome(alleles);
    genome.initializer(GARealGenome::UniformInitializer);
    genome.mutator(GARealGaussianMutator);
    genome.crossover(GARealUniformCrossover);
    GASteadyStateGA ga(genome);
    GASigmaTruncationScaling scaling;
    ga.minimize();
    GAPopulation popu = ga.population();
    popu.evaluator(pipe_evaluator);
    ga.population(popu);
    ga.populationSize(popsize);
    ga.nGenerations(generations);
    ga.pMutation(pmut);
    ga.pCrossover(pcross);
    ga.scaling(scaling);
    ga.scoreFrequency(0);
    ga.selectScores(GAStatistics::AllScores);
    ga.flushFrequency(0);

    ga.initialize(seed);
    while (!ga.done()) ga.step();

    std::printf("DONE %d\n", ga.statistics().generation());
    const GAPopulation& pop = ga.population();
    for (int i = 0; i < pop.size(); i++)
    {
        print_genes(pop.individual(i));
        std::printf(" %a\n", pop.individual(i).score());
    }
    std::printf("BEST ");
    print_genes(ga.statistics().bestIndividual());
    std::printf(" %a\n", ga.statistics().bestIndividual().score());
    std::fflush(stdout);
    return 0;
}
