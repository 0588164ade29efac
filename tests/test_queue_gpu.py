"""Work-queue campaign kernel against per-configuration launches (GPU).

Every history of the queue runs the same step with the same stream as the
single-configuration kernel. Counts (exit codes, rows) and the energy tallies,
which sum energies quantised to 2^-12 keV and so are exact in any order, agree
bit for bit with a per-history reference; a captured hipGraph replays to the
same tallies.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases():
    from pyrite_amd.queue import QueueCase

    return [
        QueueCase(Ne=300, thickness_ang=2.0e4, E0_keV=1000.0, material="W", tilt_deg=1.0, seed=3),
        QueueCase(Ne=50, thickness_ang=1.0e5, E0_keV=30000.0, material="graphite", tilt_deg=0.0, seed=4),
        QueueCase(Ne=700, thickness_ang=5.0e4, E0_keV=5000.0, material="Si", tilt_deg=2.5, seed=5),
        QueueCase(Ne=1, thickness_ang=1.0e4, E0_keV=12000.0, material="Ge", tilt_deg=0.3, seed=6),
        QueueCase(Ne=450, thickness_ang=3.0e4, E0_keV=2000.0, material="Ge", tilt_deg=3.0, seed=7),
        QueueCase(Ne=200, thickness_ang=1.5e5, E0_keV=20000.0, material="W", tilt_deg=0.5, seed=8),
    ]


def _per_config_tally(q, device):
    import torch

    from pyrite_amd import _lib
    from pyrite_amd.spectrum.geometry import tilted_geometry
    from pyrite_amd.transport.core import DeviceTables, TransportControl, make_hip_transport_core
    from pyrite_amd.transport.tables import build_layer_tables, build_transport_lut

    layers = build_layer_tables([(0.0, q.thickness_ang, list(q.composition))])
    lut = build_transport_lut(q.E_cut_keV, q.E0_keV, layers)
    core = make_hip_transport_core(DeviceTables(layers, lut, device),
                                   TransportControl(energy_model=1, straggle_on=q.straggling))
    n = q.Ne
    keys = torch.empty(n, dtype=torch.uint64, device=device)
    _lib.call("hr_stream_keys", q.seed, 0, n, keys.data_ptr(), None)
    bd, _ = tilted_geometry(np.pi / 2, np.deg2rad(q.tilt_deg))
    pos = torch.zeros((n, 3), dtype=torch.float64, device=device)
    dirs = torch.as_tensor(np.tile(bd, (n, 1)), device=device)
    E = torch.full((n,), q.E0_keV, dtype=torch.float64, device=device)
    clock = torch.zeros(n, dtype=torch.float64, device=device)
    E_cut = torch.full((n,), q.E_cut_keV, dtype=torch.float64, device=device)
    alive = torch.ones(n, dtype=torch.uint8, device=device)
    stragg = torch.zeros(n, dtype=torch.float64, device=device)
    cnt = torch.zeros(n, dtype=torch.int64, device=device)
    code = torch.zeros(n, dtype=torch.int8, device=device)
    core((0, n, 0, keys), (alive, clock, pos, dirs, E, E_cut), {}, (cnt, code), (q.straggling, stragg))
    cnt, code, Ef = cnt.cpu().numpy(), code.cpu().numpy(), E.cpu().numpy()
    t = np.zeros(12)
    for c in range(7):
        t[c] = (code == c).sum()
    t[7], t[8] = cnt.sum(), (cnt.astype(float) ** 2).sum()
    # energies quantised to 2^-12 keV as the queue tallies them: sums of the
    # quanta are exact, so the tallies agree bit for bit whatever the order
    quant = np.rint(Ef * 4096.0) / 4096.0
    t[9] = quant[code == 2].sum()
    t[10] = quant[code == 1].sum()
    esc = (code == 1) | (code == 2) | (code == 3)
    t[11] = (np.rint(q.E0_keV * 4096.0) / 4096.0 - np.where(esc, quant, 0.0)).sum()
    return t


def test_queue_matches_per_configuration_launches(device):
    from pyrite_amd.queue import TransportQueue

    cases = _cases()
    tq = TransportQueue(cases, device)
    for block, claim in ((256, 1), (64, 4), (128, 3)):
        got = tq.run(block_size=block, per_claim=claim).cpu().numpy()
        for i, q in enumerate(cases):
            want = _per_config_tally(q, device)
            np.testing.assert_array_equal(got[i], want, err_msg=f"config {i} block {block} claim {claim}")


def test_queue_graph_replays(device):
    import torch

    from pyrite_amd.queue import TransportQueue

    tq = TransportQueue(_cases(), device)
    direct = tq.run().clone()
    g = tq.graph()
    for _ in range(3):
        tq.tally.fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tq.tally.cpu().numpy(), direct.cpu().numpy())
    g.close()


def test_autotuner_picks_a_timed_candidate(device, tmp_path):
    from pyrite_amd.queue import TransportQueue

    tq = TransportQueue(_cases(), device)
    rec = tq.autotune(fraction=0.5, blocks=(64, 256), claims=(1, 4), cache=tmp_path / "tune.json")
    assert tuple(rec["best"]) in {(64, 1), (64, 4), (256, 1), (256, 4)}
    assert (tq.block_size, tq.per_claim) == tuple(rec["best"])
    rec2 = TransportQueue(_cases(), device).autotune(fraction=0.5, blocks=(64, 256), claims=(1, 4),
                                                     cache=tmp_path / "tune.json")
    assert rec2["best"] == rec["best"]
