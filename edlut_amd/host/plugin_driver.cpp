// plugin_driver.cpp -- drives one C++ plug-in (edlut_plugin.hpp) the way
// EDLUT's simulation loop does: per step, the input spikes of the step in
// delivery order (ProcessInputSpike), then UpdateState, then the spiking
// indices.  Used by the tests and as a usage example of the C++ layer.
//
//   plugin_driver selftest
//       parameter handling only (no device): defaults, SetParameters
//       errors, integration-method names; prints "ok".
//   plugin_driver run <model> <integrator> <bdf_order|n_steps> <in.bin> <out.bin>
//       in.bin : int32 n_neurons, int32 n_steps, float32 ext_current[n],
//                then per step: int32 count, int32 target[count],
//                int32 type[count], float32 weight[count]
//       out.bin: int64 n_spikes, (int32 step, int32 index)[n_spikes],
//                int32 n_vars, float32 state[n_vars][n_neurons]
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "edlut_plugin.hpp"

using namespace edlut_amd;

static std::unique_ptr<TimeDrivenNeuronModel_GPU> make(const std::string &name) {
    if (name == "AdExTimeDrivenModel_GPU") return std::make_unique<AdExTimeDrivenModel_GPU>();
    if (name == "LIFTimeDrivenModel_GPU") return std::make_unique<LIFTimeDrivenModel_GPU>();
    if (name == "HHTimeDrivenModel_GPU") return std::make_unique<HHTimeDrivenModel_GPU>();
    if (name == "EgidioGranuleCell_TimeDriven_GPU") return std::make_unique<EgidioGranuleCell_TimeDriven_GPU>();
    if (name == "TimeDrivenPurkinjeCell_GPU") return std::make_unique<TimeDrivenPurkinjeCell_GPU>();
    if (name == "IzhikevichTimeDrivenModel_GPU") return std::make_unique<IzhikevichTimeDrivenModel_GPU>();
    if (name == "LIFTimeDrivenModel_IS_GPU") return std::make_unique<LIFTimeDrivenModel_IS_GPU>();
    throw EDLUTException(EDLUT_E_ARG, "unknown neuron model " + name);
}

static int selftest() {
    AdExTimeDrivenModel_GPU m;
    ParamMap d = m.GetDefaultParameters();
    if (d.size() != 16 || std::any_cast<float>(d["g_leak"]) != 10.0f) return 1;
    bool threw = false;
    try {
        m.SetParameters({{"not_a_parameter", 1.0f}});
    } catch (const EDLUTException &e) {
        threw = e.code() == EDLUT_E_ARG;
    }
    if (!threw) return 2;
    threw = false;
    try {
        ModelDescription bad;
        bad.name = "Midpoint";
        m.SetParameters({{"int_meth", bad}});
    } catch (const EDLUTException &) {
        threw = true;
    }
    if (!threw) return 3;
    threw = false;
    try {
        m.SetParameters({{"int_meth", ModelDescription{"Bifixed_RK4", 1e-4f, 1, 0}}});
    } catch (const EDLUTException &e) {
        threw = e.code() == EDLUT_E_ARG;
    }
    if (!threw) return 8;
    m.SetParameters({{"int_meth", ModelDescription{"Bifixed_Euler", 1e-4f, 1, 3}}});
    m.SetParameters({{"g_leak", 12.5}, {"int_meth", ModelDescription{"RK4", 1e-4f, 1}}});
    ParamMap p = m.GetParameters();
    if (std::any_cast<float>(p["g_leak"]) != 12.5f) return 4;
    if (std::any_cast<ModelDescription>(p["int_meth"]).name != "RK4") return 5;
    if (m.CheckSynapseType(3, true) || !m.CheckSynapseType(0, true) || !m.CheckSynapseType(3, false)) return 6;
    threw = false;
    try {
        (void)m.GetVectorNeuronState(0);
    } catch (const EDLUTException &e) {
        threw = e.code() == EDLUT_E_STATE;
    }
    if (!threw) return 7;
    std::printf("ok\n");
    return 0;
}

template <typename T> static bool rd(FILE *f, T *p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }

static int run(int argc, char **argv) {
    if (argc < 7) return 64;
    auto m = make(argv[2]);
    const float dt = 1e-4f;
    const int iparam = std::atoi(argv[4]);   // BDF order, or the Bifixed methods' n_steps
    m->SetParameters({{"int_meth", ModelDescription{argv[3], dt, iparam, iparam}}});
    FILE *in = std::fopen(argv[5], "rb");
    if (!in) return 66;
    int32_t n = 0, steps = 0;
    if (!rd(in, &n, 1) || !rd(in, &steps, 1)) return 65;
    std::vector<float> ext((size_t)n);
    if (!rd(in, ext.data(), ext.size())) return 65;
    for (int t = 0; t < 3; t++) m->CheckSynapseType(t, true);
    m->CheckSynapseType(3, false);
    m->InitializeStates(n, 0, 0);
    for (int i = 0; i < n; i++) m->ProcessInputCurrent(i, ext[(size_t)i]);
    std::vector<int32_t> out_step, out_idx;
    double t_now = 0.0;
    for (int k = 1; k <= steps; k++) {
        int32_t c = 0;
        if (!rd(in, &c, 1)) return 65;
        std::vector<int32_t> tg((size_t)c), ty((size_t)c);
        std::vector<float> w((size_t)c);
        if (c && (!rd(in, tg.data(), (size_t)c) || !rd(in, ty.data(), (size_t)c) || !rd(in, w.data(), (size_t)c)))
            return 65;
        for (int32_t e = 0; e < c; e++) m->ProcessInputSpike(tg[(size_t)e], ty[(size_t)e], w[(size_t)e]);
        // update times accumulate double(float(dt)) (TimeEventUpdateNeurons.cpp:78)
        t_now += (double)dt;
        m->UpdateState(-1, t_now);
        for (int32_t i : m->GetInternalSpikeIndexs()) {
            out_step.push_back(k);
            out_idx.push_back(i);
        }
    }
    std::fclose(in);
    FILE *out = std::fopen(argv[6], "wb");
    if (!out) return 73;
    const int64_t ns = (int64_t)out_step.size();
    std::fwrite(&ns, sizeof(ns), 1, out);
    for (int64_t q = 0; q < ns; q++) {
        std::fwrite(&out_step[(size_t)q], 4, 1, out);
        std::fwrite(&out_idx[(size_t)q], 4, 1, out);
    }
    int nv = 0, nd = 0;
    int64_t stride = 0;
    check(edlut_population_layout(m->handle(), &nv, &nd, &stride));
    const int32_t nv32 = nv;
    std::fwrite(&nv32, 4, 1, out);
    for (int v = 0; v < nv; v++) {
        std::vector<float> s = m->GetVectorNeuronState(v);
        std::fwrite(s.data(), sizeof(float), s.size(), out);
    }
    std::fclose(out);
    return 0;
}

int main(int argc, char **argv) {
    try {
        if (argc >= 2 && std::strcmp(argv[1], "selftest") == 0) return selftest();
        if (argc >= 2 && std::strcmp(argv[1], "run") == 0) return run(argc, argv);
    } catch (const EDLUTException &e) {
        std::fprintf(stderr, "EDLUTException: %s\n", e.what());
        return 70;
    }
    std::fprintf(stderr, "usage: plugin_driver selftest | run <model> <integrator> <bdf_order|n_steps> <in> <out>\n");
    return 64;
}
