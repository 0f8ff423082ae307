/*
 * ref_dump.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Driver linked against two of the reference's own source files,
 * Kernel/src/simulation/ExponentialTable.cpp and
 * Kernel/src/simulation/RandomGenerator.cpp, compiled where they lie (see
 * oracle/Makefile, target `ref`).  It writes what those files compute so that
 * the oracle's restatement can be checked against the reference itself:
 *
 *   <dir>/exp_table.f32      ExponentialTable::LookUpTable, 16384 floats
 *   <dir>/exp_probe.f32      ExponentialTable::GetResult on the probe inputs
 *   <dir>/lcg_seeds.u32      states of 64 generators built after srand(12345)
 *   <dir>/lcg_frand.f32      4096 frand() draws of the first generator
 *
 * The probe inputs are read from <dir>/exp_probe_in.f32 (written by the test).
 * `using namespace std` precedes the include, as in the reference's model
 * translation units (Kernel/include/neuron_model/NeuronModel.h), so that the
 * inline GetResult resolves exp() the way the models see it.
 */
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include <iostream>
using namespace std;
#include "simulation/ExponentialTable.h"
#include "simulation/RandomGenerator.h"

static void write_file(const std::string &path, const void *data, size_t bytes) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(2); }
    fwrite(data, 1, bytes, f);
    fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 2) { fprintf(stderr, "usage: ref_dump <dir>\n"); return 2; }
    std::string dir = argv[1];
    write_file(dir + "/exp_table.f32", ExponentialTable::LookUpTable,
               sizeof(float) * ExponentialTable::TableSize);

    std::vector<float> in;
    FILE *f = fopen((dir + "/exp_probe_in.f32").c_str(), "rb");
    if (f) {
        float x;
        while (fread(&x, sizeof(float), 1, f) == 1) in.push_back(x);
        fclose(f);
    }
    std::vector<float> out(in.size());
    for (size_t i = 0; i < in.size(); i++) out[i] = ExponentialTable::GetResult(in[i]);
    write_file(dir + "/exp_probe.f32", out.data(), sizeof(float) * out.size());

    RandomGenerator::srand(12345);
    std::vector<unsigned int> seeds;
    std::vector<RandomGenerator *> gens;
    for (int i = 0; i < 64; i++) {
        gens.push_back(new RandomGenerator());
        seeds.push_back((unsigned int)gens.back()->next_element);
    }
    write_file(dir + "/lcg_seeds.u32", seeds.data(), sizeof(unsigned int) * seeds.size());
    std::vector<float> fr(4096);
    for (int i = 0; i < 4096; i++) fr[i] = gens[0]->frand();
    write_file(dir + "/lcg_frand.f32", fr.data(), sizeof(float) * fr.size());
    for (RandomGenerator *g : gens) delete g;
    return 0;
}
