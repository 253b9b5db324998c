"""One small forward solve on cuda:0 checked against the CPU oracle."""
import numpy as np


def run_smoke():
    import torch
    from oracle import oracle as orc
    from radiativetransfermontecarlo_amd import solver
    from radiativetransfermontecarlo_amd.material import MaterialClass
    from radiativetransfermontecarlo_amd.observation import initializeObservation
    from tests import cases
    from tests.stats import assert_same_distribution

    if not torch.cuda.is_available():
        raise RuntimeError("smoke needs a GPU")
    geometry, source, material, observation = cases.config1(n=200_000)
    observation = dict(observation, time=np.arange(0, 1 + 1e-9, 0.02))
    obs = solver.radiativeTransfer(geometry, source, material, observation, seed=1, device="cuda:0",
                                   return_histogram=True)
    assert np.all(obs['energy'][1:] == 1.0), "energy not conserved at Q = Inf"
    O = initializeObservation(geometry, True, observation, 20_000)
    mat = MaterialClass.prepareSigma(cases.config1()[2], 3)
    Eo, _ = orc.forward(geometry, source, mat, O, 20_000, seed=2, packet=10_000)
    res = assert_same_distribution(obs['E'].reshape(-1, order='F'), obs['N'], Eo, 20_000, "smoke")
    print(f"smoke ok: {res}")


if __name__ == "__main__":
    run_smoke()
