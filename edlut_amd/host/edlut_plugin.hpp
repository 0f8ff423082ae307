// edlut_plugin.hpp -- C++ host plug-ins over the C ABI (include/edlut_amd.h).
//
// Mirrors EDLUT's GPU neuron-model plug-in interface
// (TimeDrivenNeuronModel_GPU_C_INTERFACE, Kernel/include/neuron_model/
// TimeDrivenNeuronModel_GPU_C_INTERFACE.cuh:64-423) with the same method
// names, parameter names, defaults and error behaviour, so that a host
// simulator written against those plug-ins can drive the gfx950 kernels:
//
//   SetParameters / GetParameters / GetDefaultParameters  (std::any instead
//       of boost::any; unknown keys throw EDLUTException)
//   InitializeStates(N, queue)        allocates the device population
//   CheckSynapseType(type, spiking)   enables a synapse type
//   ProcessInputSpike(target, type, weight)   queued in call order
//   ProcessInputCurrent(target, total_current)
//   UpdateState(index, CurrentTime)   one synchronous step; returns false
//   GetInternalSpikeIndexs()          ascending spiking indices of the step
//   GetVectorNeuronState(var)         copy of one state variable
//
// The reference passes Interconnection / Neuron objects; this layer takes
// the fields it reads from them (target index, synapse type, weight).
// Header-only; link with -ledlut_amd.
#pragma once

#include <any>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/edlut_amd.h"

namespace edlut_amd {

// EDLUTException (Kernel/src/spike/EDLUTException.cpp): configuration and
// device errors surface as exceptions carrying the library's message.
class EDLUTException : public std::runtime_error {
  public:
    EDLUTException(int code, const std::string &msg)
        : std::runtime_error("[" + std::to_string(code) + "] " + msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc) {
    if (rc != 0) throw EDLUTException(rc, edlut_last_error());
}

using ParamMap = std::map<std::string, std::any>;

// Integration-method description (IntegrationMethod_GPU_C_INTERFACE: "name",
// "step" in seconds, "bdf_order" for BDF, BDFn_GPU_C_INTERFACE.cuh:275-288;
// "n_steps" for the Bifixed methods, BifixedStep_GPU_C_INTERFACE.cuh:144-171).
struct ModelDescription {
    std::string name = "Euler";
    float step = 1e-3f;
    int bdf_order = 1;
    int n_steps = 2;
};

inline int integrator_id(const std::string &name) {
    if (name == "Euler") return EDLUT_EULER;
    if (name == "RK2") return EDLUT_RK2;
    if (name == "RK4") return EDLUT_RK4;
    if (name == "BDF") return EDLUT_BDF;
    if (name == "Bifixed_Euler") return EDLUT_BIFIXED_EULER;
    if (name == "Bifixed_RK2") return EDLUT_BIFIXED_RK2;
    if (name == "Bifixed_RK4") return EDLUT_BIFIXED_RK4;
    // IntegrationMethodFactory_GPU_C_INTERFACE.cu:107-110
    throw EDLUTException(EDLUT_E_ARG, "unknown integration method " + name);
}

inline float as_float(const std::any &v, const std::string &key) {
    if (v.type() == typeid(float)) return std::any_cast<float>(v);
    if (v.type() == typeid(double)) return (float)std::any_cast<double>(v);
    if (v.type() == typeid(int)) return (float)std::any_cast<int>(v);
    throw EDLUTException(EDLUT_E_ARG, "parameter " + key + " must be a number");
}

// Common part of the GPU time-driven plug-ins.
class TimeDrivenNeuronModel_GPU {
  public:
    virtual ~TimeDrivenNeuronModel_GPU() {
        if (pop_) edlut_population_destroy(pop_);
    }
    TimeDrivenNeuronModel_GPU(const TimeDrivenNeuronModel_GPU &) = delete;
    TimeDrivenNeuronModel_GPU &operator=(const TimeDrivenNeuronModel_GPU &) = delete;

    virtual std::string GetName() const = 0;

    // GetDefaultParameters: every model parameter, the integration method
    // and the model name.
    ParamMap GetDefaultParameters() const {
        ParamMap m;
        for (const auto &kv : defaults_) m[kv.first] = kv.second;
        m["int_meth"] = ModelDescription{};
        m["name"] = GetName();
        return m;
    }
    ParamMap GetParameters() const {
        ParamMap m;
        for (size_t i = 0; i < defaults_.size(); i++) m[defaults_[i].first] = values_[i];
        m["int_meth"] = int_meth_;
        m["name"] = GetName();
        return m;
    }
    // SetParameters: unknown keys and a changed model name throw; values are
    // fixed once the states are initialised.
    void SetParameters(const ParamMap &params) {
        if (pop_) throw EDLUTException(EDLUT_E_STATE, "parameters are fixed once the states are initialised");
        for (const auto &kv : params) {
            if (kv.first == "name") {
                if (std::any_cast<std::string>(kv.second) != GetName())
                    throw EDLUTException(EDLUT_E_ARG, "model name does not match " + GetName());
            } else if (kv.first == "int_meth") {
                ModelDescription d = std::any_cast<ModelDescription>(kv.second);
                const int id = integrator_id(d.name);
                if (!(d.step > 0.0f)) throw EDLUTException(EDLUT_E_ARG, "integration step must be positive");
                if (id >= EDLUT_BIFIXED_EULER && d.n_steps <= 0)
                    throw EDLUTException(EDLUT_E_ARG, "n_steps must be positive");
                int_meth_ = d;
            } else {
                size_t i = 0;
                while (i < defaults_.size() && defaults_[i].first != kv.first) i++;
                if (i == defaults_.size())
                    throw EDLUTException(EDLUT_E_ARG, "unknown parameter " + kv.first + " for " + GetName());
                values_[i] = as_float(kv.second, kv.first);
            }
        }
    }

    void InitializeStates(int N_neurons, int /*OpenMPQueueIndex*/ = 0, int device = 0) {
        if (pop_) throw EDLUTException(EDLUT_E_STATE, "states already initialised");
        const int integ = integrator_id(int_meth_.name);
        check(edlut_population_create(model_, integ, integ >= EDLUT_BIFIXED_EULER ? int_meth_.n_steps : int_meth_.bdf_order,
                                      values_.data(), (int)values_.size(), int_meth_.step, N_neurons,
                                      device, &pop_));
        n_ = N_neurons;
        if (mask_) check(edlut_population_enable_synapses(pop_, mask_));
    }

    // CheckSynapseType (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:239-290): types
    // 0-2 (AMPA, GABA, NMDA) need spiking sources, type 3 a current source.
    bool CheckSynapseType(int type, bool source_emits_spikes) {
        if (type < 0 || type > 3 || (type < 3) != source_emits_spikes) return false;
        mask_ |= 1u << type;
        if (pop_) check(edlut_population_enable_synapses(pop_, 1u << type));
        return true;
    }

    // ProcessInputSpike (AdExTimeDrivenModel_GPU_C_INTERFACE.cu:69-73): the
    // weight is added to the target's conductance before the next update, in
    // the order of the calls.
    void ProcessInputSpike(int target, int type, float weight) {
        tgt_.push_back(target);
        type_.push_back((uint8_t)type);
        w_.push_back(weight);
    }

    // ProcessInputCurrent (:75-82) with the total current of the target.
    void ProcessInputCurrent(int target, float total_current) {
        need();
        if (ext_.empty()) ext_.assign((size_t)n_, 0.0f);
        ext_.at((size_t)target) = total_current;
        ext_dirty_ = true;
    }

    // UpdateState (:90-112): one synchronous step; spikes are read with
    // GetInternalSpikeIndexs().  The update index is CurrentTime / step.
    bool UpdateState(int /*index*/, double CurrentTime) {
        need();
        if (ext_dirty_) {
            check(edlut_population_set_ext_current(pop_, ext_.data()));
            ext_dirty_ = false;
        }
        if (!tgt_.empty())
            check(edlut_population_push_inputs(pop_, tgt_.data(), type_.data(), w_.data(), (int)tgt_.size()));
        tgt_.clear();
        type_.clear();
        w_.clear();
        const int64_t step = (int64_t)std::llround(CurrentTime / (double)int_meth_.step);
        int ns = 0;
        check(edlut_population_update(pop_, step, &ns));
        spikes_.resize((size_t)ns);
        if (ns) check(edlut_population_get_spikes(pop_, spikes_.data(), ns));
        return false;
    }

    const std::vector<int32_t> &GetInternalSpikeIndexs() const { return spikes_; }

    std::vector<float> GetVectorNeuronState(int var) const {
        need();
        std::vector<float> out((size_t)n_);
        check(edlut_population_get_state(pop_, var, out.data()));
        return out;
    }

    int GetSize() const { return n_; }
    edlut_population *handle() const { return pop_; }

    // TimeDrivenNeuronModel_GPU_C_INTERFACE.cuh / NeuronModel.h identifiers
    std::string GetModelSimulationMethod() const { return "TIME_DRIVEN_MODEL_GPU"; }
    std::string GetModelType() const { return "NEURAL_LAYER"; }
    std::string GetModelOutputActivityType() const { return "OUTPUT_SPIKE"; }
    std::string GetModelInputActivityType() const { return "INPUT_SPIKE_AND_CURRENT"; }

  protected:
    TimeDrivenNeuronModel_GPU(int model, std::vector<std::pair<std::string, float>> defaults)
        : model_(model), defaults_(std::move(defaults)) {
        for (const auto &kv : defaults_) values_.push_back(kv.second);
    }

  private:
    void need() const {
        if (!pop_) throw EDLUTException(EDLUT_E_STATE, "InitializeStates has not been called");
    }
    int model_;
    std::vector<std::pair<std::string, float>> defaults_;
    std::vector<float> values_;
    ModelDescription int_meth_{};
    edlut_population *pop_ = nullptr;
    int n_ = 0;
    uint32_t mask_ = 0;
    std::vector<int32_t> tgt_;
    std::vector<uint8_t> type_;
    std::vector<float> w_;
    std::vector<float> ext_;
    bool ext_dirty_ = false;
    std::vector<int32_t> spikes_;
};

// AdExTimeDrivenModel_GPU (defaults of AdExTimeDrivenModel.cpp:449-467).
class AdExTimeDrivenModel_GPU : public TimeDrivenNeuronModel_GPU {
  public:
    AdExTimeDrivenModel_GPU()
        : TimeDrivenNeuronModel_GPU(EDLUT_ADEX, {{"a", 1.0f}, {"b", 9.0f}, {"thr_slo_fac", 2.0f},
                                                 {"v_thr", -50.0f}, {"tau_w", 50.0f}, {"e_exc", 0.0f},
                                                 {"e_inh", -80.0f}, {"e_reset", -80.0f}, {"e_leak", -65.0f},
                                                 {"g_leak", 10.0f}, {"c_m", 110.0f}, {"tau_exc", 5.0f},
                                                 {"tau_inh", 10.0f}, {"tau_nmda", 20.0f}}) {}
    std::string GetName() const override { return "AdExTimeDrivenModel_GPU"; }
};

// LIFTimeDrivenModel_GPU (defaults of LIFTimeDrivenModel.cpp:435-450).
class LIFTimeDrivenModel_GPU : public TimeDrivenNeuronModel_GPU {
  public:
    LIFTimeDrivenModel_GPU()
        : TimeDrivenNeuronModel_GPU(EDLUT_LIF, {{"e_exc", 0.0f}, {"e_inh", -80.0f}, {"e_leak", -65.0f},
                                                {"v_thr", -50.0f}, {"c_m", 2.0f}, {"tau_exc", 5.0f},
                                                {"tau_inh", 10.0f}, {"tau_ref", 1.0f}, {"g_leak", 0.2f},
                                                {"tau_nmda", 20.0f}}) {}
    std::string GetName() const override { return "LIFTimeDrivenModel_GPU"; }
};

// HHTimeDrivenModel_GPU (defaults of HHTimeDrivenModel.cpp GetDefaultParameters).
class HHTimeDrivenModel_GPU : public TimeDrivenNeuronModel_GPU {
  public:
    HHTimeDrivenModel_GPU()
        : TimeDrivenNeuronModel_GPU(EDLUT_HH, {{"e_exc", 0.0f}, {"e_inh", -80.0f}, {"e_leak", -65.0f},
                                               {"g_leak", 10.0f}, {"c_m", 120.0f}, {"v_thr", -52.0f},
                                               {"tau_exc", 5.0f}, {"tau_inh", 10.0f}, {"tau_nmda", 20.0f},
                                               {"g_na", 20000.0f}, {"g_kd", 6000.0f}, {"e_na", 50.0f},
                                               {"e_k", -90.0f}}) {}
    std::string GetName() const override { return "HHTimeDrivenModel_GPU"; }
    void SetNewtonFP64(bool on) { check(edlut_population_set_newton_fp64(handle(), on ? 1 : 0)); }
};

// EgidioGranuleCell_TimeDriven_GPU: fixed constants, no parameters besides
// the integration method (EgidioGranuleCell_TimeDriven.cpp:26-58).
class EgidioGranuleCell_TimeDriven_GPU : public TimeDrivenNeuronModel_GPU {
  public:
    EgidioGranuleCell_TimeDriven_GPU() : TimeDrivenNeuronModel_GPU(EDLUT_GRANULE, {}) {}
    std::string GetName() const override { return "EgidioGranuleCell_TimeDriven_GPU"; }
};

// TimeDrivenPurkinjeCell_GPU (defaults of TimeDrivenPurkinjeCell.cpp:423-434).
class TimeDrivenPurkinjeCell_GPU : public TimeDrivenNeuronModel_GPU {
  public:
    TimeDrivenPurkinjeCell_GPU()
        : TimeDrivenNeuronModel_GPU(EDLUT_PURKINJE, {{"e_exc", 0.0f}, {"e_inh", -80.0f}, {"v_thr", -35.0f},
                                                     {"e_leak", -70.0f}, {"tau_exc", 1.0f}, {"tau_inh", 2.0f},
                                                     {"tau_nmda", 20.0f}, {"tau_ref", 1.35f}}) {}
    std::string GetName() const override { return "TimeDrivenPurkinjeCell_GPU"; }
};

// IzhikevichTimeDrivenModel_GPU (defaults of IzhikevichTimeDrivenModel.cpp:389-401).
class IzhikevichTimeDrivenModel_GPU : public TimeDrivenNeuronModel_GPU {
  public:
    IzhikevichTimeDrivenModel_GPU()
        : TimeDrivenNeuronModel_GPU(EDLUT_IZHIKEVICH, {{"a", 0.1f}, {"b", 0.23f}, {"c", -65.0f}, {"d", 0.2f},
                                                       {"e_exc", 0.0f}, {"e_inh", -80.0f}, {"c_m", 100.0f},
                                                       {"tau_exc", 5.0f}, {"tau_inh", 10.0f},
                                                       {"tau_nmda", 20.0f}}) {}
    std::string GetName() const override { return "IzhikevichTimeDrivenModel_GPU"; }
};

// LIFTimeDrivenModel_IS_GPU: SI units (defaults of LIFTimeDrivenModel_IS.cpp:392-404).
class LIFTimeDrivenModel_IS_GPU : public TimeDrivenNeuronModel_GPU {
  public:
    LIFTimeDrivenModel_IS_GPU()
        : TimeDrivenNeuronModel_GPU(EDLUT_LIF_IS, {{"e_exc", 0.0f}, {"e_inh", -80.0e-3f}, {"e_leak", -65.0e-3f},
                                                   {"v_thr", -50.0e-3f}, {"c_m", 2.0e-12f}, {"tau_exc", 5.0e-3f},
                                                   {"tau_inh", 10.0e-3f}, {"tau_ref", 1.0e-3f},
                                                   {"g_leak", 0.2e-9f}, {"tau_nmda", 20.0e-3f}}) {}
    std::string GetName() const override { return "LIFTimeDrivenModel_IS_GPU"; }
};

}  // namespace edlut_amd
