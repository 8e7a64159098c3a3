"""GPU tests of captured launches (include/dengo_graph.h): a graph replayed
several times, with fresh inputs copied into its buffers and the program's
own eager work and synchronisations between replays, gives every time the
bit-identical result of the eager evolve_device call."""
import os

import numpy as np
import pytest
import torch

from dengo_amd import runtime
from dengo_amd.ics import oxygen_ics, primordial_ics, tolerances
from dengo_amd.synthetic import SyntheticNetwork

pytestmark = pytest.mark.gpu
GOLD_SYN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "synthetic_cells.npz")


def _eager(solver, x, rt, at, dtf, prec, T=None):
    dev = torch.device("cuda:0")
    dx, drt, dat = (torch.from_numpy(a).to(dev) for a in (x, rt, at))
    st = torch.zeros(x.shape[0], dtype=torch.int32, device=dev)
    if T is None:
        solver.evolve_device(dx, drt, dat, dtf, outputs={"status": st}, precision=prec)
    else:
        solver.evolve_device(dx, torch.from_numpy(T).to(dev), drt, dat, dtf,
                             outputs={"status": st}, precision=prec)
    return dx.cpu().numpy(), dat.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("net", ["primordial", "oxygen", "synthetic"])
def test_graph_replays_match_eager(net):
    dev = torch.device("cuda:0")
    T = None
    if net == "primordial":
        solver = runtime.PrimordialSolver()
        inputs = [primordial_ics(700, seed=s)[0] for s in (1, 2, 3)]
        dtf, prec = 1e11, runtime.FP64_FAST
    elif net == "oxygen":
        solver = runtime.OxygenSolver()
        inputs = [oxygen_ics(700, seed=s)[0] for s in (4, 5, 6)]
        dtf, prec = 1e11, runtime.MIXED
    else:
        solver = runtime.SyntheticSolver(SyntheticNetwork(seed=7))
        g = np.load(GOLD_SYN)
        T = np.ascontiguousarray(g["T"][:6])
        inputs = [np.ascontiguousarray(g["input"][:6] * f) for f in (1.0, 1.1, 0.9)]
        dtf, prec = float(g["dtf"]), runtime.FP64_FAST
    try:
        n = inputs[0].shape[0]
        rts = [tolerances(x) for x in inputs]
        bx = torch.empty((n, solver.N), dtype=torch.float64, device=dev)
        brt = torch.empty_like(bx)
        bat = torch.empty_like(bx)
        bst = torch.zeros(n, dtype=torch.int32, device=dev)
        dT = None if T is None else torch.from_numpy(T).to(dev)
        cap = runtime.CapturedEvolve(solver, bx, brt, bat, dtf, T=dT, outputs={"status": bst},
                                     precision=prec)
        stream = torch.cuda.Stream()
        for rep in range(2):
            for x, (rt, at) in zip(inputs, rts):
                bx.copy_(torch.from_numpy(x))
                brt.copy_(torch.from_numpy(rt))
                bat.copy_(torch.from_numpy(at))
                bst.zero_()
                torch.cuda.synchronize()
                cap.launch(stream)
                stream.synchronize()
                want = _eager(solver, x, rt, at, dtf, prec, T)
                assert np.array_equal(bx.cpu().numpy(), want[0]), (net, rep)
                assert np.array_equal(bat.cpu().numpy(), want[1]), (net, rep)
                assert np.array_equal(bst.cpu().numpy(), want[2]), (net, rep)
        cap.close()
    finally:
        solver.close()
