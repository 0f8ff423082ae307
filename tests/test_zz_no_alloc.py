"""The allocation-free update (edlut_population_update_noalloc) against the
ordinary one (edlut_population_push_inputs + edlut_population_update): the
same spike lists and bit-identical state after every update, on calls in a
row with fewer events each time, events and spikes in device memory."""
import numpy as np
import pytest
import torch

import edlut_amd as E
from edlut_amd import _capi as C
from tests.helpers import bits

pytestmark = pytest.mark.gpu


CASES = [("AdExTimeDrivenModel_GPU", "RK4", 12.0),
         # the granule's BDF kernel: lanes refilled from equal ranges, spike lists
         # written by the last workgroup
         ("EgidioGranuleCell_TimeDriven_GPU", "BDF", 1.0)]


def _pop(n, model="AdExTimeDrivenModel_GPU", integ="RK4"):
    m = E.CreateNeuronModel(model, {"int_meth": E.integration_method(integ, 1e-4, 2 if integ == "BDF" else 1)})
    for t in (0, 1, 2):
        assert m.CheckSynapseType(t)
    m.InitializeStates(n, 0, 0)
    return m


@pytest.mark.parametrize("model,integ,wmax", CASES)
def test_update_noalloc_equals_update(model, integ, wmax):
    n = 4000
    print(f"case noalloc update {model} {integ} n=4000", flush=True)
    ref, m = _pop(n, model, integ), _pop(n, model, integ)
    rng = np.random.default_rng(11)
    idx = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch_stream = torch.cuda.current_stream().cuda_stream
    fired = 0
    for k, ne in enumerate((300_000, 40_000, 200_000, 1_000, 0, 90_000), start=1):
        tgt = rng.integers(0, n, ne).astype(np.int32)
        tgt[: ne // 10] = 7                     # one hot target
        typ = rng.choice(np.array([0, 0, 0, 1, 2], np.uint8), ne)
        w = rng.uniform(0.0, wmax, ne).astype(np.float32)
        ref.ProcessInputSpikes(tgt, typ, w)
        ref.UpdateState(step=k)
        want = ref.GetInternalSpikeIndexs()
        d_t, d_y, d_w = (torch.from_numpy(a).to("cuda") for a in (tgt, typ, w))
        torch.cuda.synchronize()
        stream = torch_stream if k % 2 else None   # torch's stream, then the population's
        C.check(C.lib().edlut_population_update_noalloc(m.handle, k, d_t.data_ptr(), d_y.data_ptr(), d_w.data_ptr(),
                                                        ne, idx.data_ptr(), cnt.data_ptr(), stream))
        torch.cuda.synchronize()
        c = int(cnt.item())
        got = idx[:c].cpu().numpy()
        assert np.array_equal(got, want), f"spikes of update {k} differ"
        fired += c
        for v in range(len(m.STATE_VARS)):
            assert np.array_equal(bits(m.GetVectorNeuronState(v)), bits(ref.GetVectorNeuronState(v))), \
                f"update {k}, state {v}"
    assert fired > 0, "the drive should make neurons fire"
    # more events than the device array holds: refused before any launch
    with pytest.raises(C.EDLUTException) as e:
        C.check(C.lib().edlut_population_update_noalloc(m.handle, 7, idx.data_ptr(), idx.data_ptr(), idx.data_ptr(),
                                                        (1 << 20) + 1, idx.data_ptr(), cnt.data_ptr(), None))
    assert e.value.code == C.EDLUT_E_CAP
    ref.close()
    m.close()
