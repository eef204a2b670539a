"""Write the oracle's golden vectors: tests/golden/golden_<config>.json with seeded inputs
(four-momenta of 64 data and 192 MC events, data weights, coupling values) and the FP64 oracle's
outputs on them (per-event |A|^2 of data and MC, NLL, gradient and autograd Hessian in
trainable_vars order). The GPU tests compare the engine with these committed numbers; a CPU test
re-runs the oracle on the same inputs so that a change of the oracle shows up as a fixture
mismatch. Run in the build container: python tests/golden/make_golden.py [case names]"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASES = {"lc_pkpi_2res": 101, "lc_pkpi_6res": 202, "jpsi_kkpipi_12ch": 303}
N_DATA, N_MC = 64, 192
# four-body cases: weighted flat phase space as data (the likelihood does not assume the data
# follow the model), from the host GENBOD generator
FOUR_BODY = {"jpsi_kkpipi_12ch"}


def config_file(name):
    return os.path.join(ROOT, "tf-pwa_amd", "tf_pwa_amd", "configs", name + ".yml")


def compute(name, inputs):
    """Oracle outputs for a fixture's inputs (also used by the CPU regression test)."""
    from oracle import tfpwa_oracle as O

    om = O.Model(config_file(name))
    om.set_params(inputs["params"])
    order = om.config["data"]["dat_order"]
    data = {k: np.array(inputs["data_p4"][k]) for k in order}
    mc = {k: np.array(inputs["mc_p4"][k]) for k in order}
    ad, am = O.cal_angle(om, data), O.cal_angle(om, mc)
    var = om.var_tensors()
    f_data = O.sum_amp(om, ad, var).numpy()
    f_mc = O.sum_amp(om, am, var).numpy()
    w, wm = O.fcn_weights(len(data[order[0]]), np.array(inputs["data_weight"]),
                          len(mc[order[0]]))
    nll, g, H = O.nll_grad_hessian(om, ad, am, w, wm)
    return {"f_data": f_data.tolist(), "f_mc": f_mc.tolist(), "nll": float(nll),
            "grad": np.asarray(g).tolist(), "hessian": np.asarray(H).tolist(),
            "trainable": list(om.trainable)}


def main(names=None):
    from oracle import tfpwa_oracle as O

    for name, seed in CASES.items():
        if names and name not in names:
            continue
        om = O.Model(config_file(name))
        params = O.random_couplings(om, seed)
        om.set_params(params)
        if name in FOUR_BODY:
            sys.path.insert(0, os.path.join(ROOT, "tf-pwa_amd"))
            from tf_pwa_amd.phasespace import generate_phsp

            order = om.config["data"]["dat_order"]
            masses = [om.particles[n].mass for n in order]
            data = dict(zip(order, generate_phsp(om.top.mass, masses, N_DATA,
                                                 np.random.default_rng(seed + 1))))
            mc = dict(zip(order, generate_phsp(om.top.mass, masses, N_MC,
                                               np.random.default_rng(seed + 2))))
        else:
            data = O.gen_data(om, N_DATA, seed + 1, batch=20000)
            mc = O.gen_mc(om, N_MC, seed + 2)
        wd = np.random.default_rng(seed + 3).uniform(0.5, 1.5, N_DATA)
        inputs = {"config": name + ".yml", "seed": seed,
                  "params": {k: float(v) for k, v in om.values.items()},
                  "data_p4": {k: np.asarray(v).tolist() for k, v in data.items()},
                  "mc_p4": {k: np.asarray(v).tolist() for k, v in mc.items()},
                  "data_weight": wd.tolist()}
        out = compute(name, inputs)
        path = os.path.join(HERE, "golden_{}.json".format(name))
        with open(path, "w") as fh:
            json.dump({"inputs": inputs, "outputs": out}, fh)
        print(name, "nll", out["nll"], "->", path)


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
