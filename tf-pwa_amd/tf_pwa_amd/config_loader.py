"""config.yml front end: the subset of tf_pwa.config_loader.ConfigLoader that builds the
likelihood (decay group, parameters, data, Model and FCN) and drives a fit.

Follows tf_pwa/config_loader/decay_config.py:49-427 (particles, decays, chain enumeration)
and config_loader.py:65-104, 210-237, 358-537 (parameters and constraints), 539-720
(model/FCN construction), 744-801 (fit) and 995-1037 (parameter IO).
"""

import json

import numpy as np
import yaml

from .amp import AmplitudeModel
from .data import EventSet
from .model import FCN, Model
from .particle import Decay, DecayChain, DecayGroup, Particle, chain_decays
from .variable import VarsManager

_PARTICLE_KEYS = {"Par": "P", "m0": "mass", "g0": "width", "bw": "model"}


def load_config(config):
    if isinstance(config, dict):
        return dict(config)
    with open(config) as f:
        return yaml.safe_load(f) or {}


def load_dat_file(fnames, n_final):
    """Whitespace-separated four-momenta, ``n_final`` rows per event
    (tf_pwa/data.py:300-364)."""
    if isinstance(fnames, str):
        fnames = [fnames]
    arrs = [np.loadtxt(f, dtype=np.float64).reshape(-1, n_final, 4) for f in fnames]
    return np.concatenate(arrs, axis=0)


def inverse_hessian(h, force_pos=True):
    """Error matrix from a Hessian: the inverse when it is positive definite; otherwise, as in
    force_pos_def (tf_pwa/applications.py:249-277), eigenvalues that are negative only by
    rounding (above -1e-5 of the largest) are lifted to positive, and larger negative
    eigenvalues are replaced by their magnitude (a positive-definite stand-in)."""
    h = 0.5 * (np.asarray(h) + np.asarray(h).T)
    e, v = np.linalg.eigh(h)
    if e.min() > 0 or not force_pos:
        return np.linalg.inv(h) if e.min() > 0 else np.linalg.pinv(h)
    emax = np.abs(e).max()
    if e.min() > -1e-5 * emax:
        e_new = np.where(e < 0, e - e.min() * 1.1, e)
    else:
        e_new = np.where(np.abs(e) > 1e-12 * emax, np.abs(e), 1e-12 * emax)
    return (v / e_new) @ v.T


class ConfigLoader:
    """Build the decay group, variables, data and likelihood from a config.yml dict/file."""

    def __init__(self, config, seed=None):
        self.config = load_config(config)
        self.rng = np.random.default_rng(seed)
        self.data_config = self.config.get("data", {}) or {}
        self._build_decay_group()
        self.vm = VarsManager(polar=self.config.get("polar", True))
        self._init_params()
        self.amp = None
        self.inmc = None

    # ------------------------------------------------------------------ decay group
    def _build_decay_group(self):
        pconf = dict(self.config["particle"])
        top = pconf.pop("$top")
        finals = pconf.pop("$finals")
        props = {}
        for block in (top, finals, pconf):
            for name, prop in block.items():
                props[name] = {_PARTICLE_KEYS.get(k, k): v for k, v in (prop or {}).items()}
        self.particles = {n: Particle(n, **p) for n, p in props.items()}
        self.top_name = list(top.keys())[0]
        self.final_names = list(finals.keys())
        for core, outs in self.config["decay"].items():
            items = outs if all(isinstance(i, list) for i in outs) else [outs]
            for it in items:
                parts = [x for x in it if not isinstance(x, dict)]
                params = {}
                for x in it:
                    if isinstance(x, dict):
                        params.update(x)
                d = Decay(self.particles[core], [self.particles[p] for p in parts], **params)
                self.particles[core].decays.append(d)
        top_p = self.particles[self.top_name]
        finals_set = sorted(self.final_names)
        chains = []
        for c in chain_decays(top_p):
            dc = DecayChain(c)
            if sorted(p.name for p in dc.outs) == finals_set:
                chains.append(dc)
        if not chains:
            raise RuntimeError("no decay chain available, check the config")
        self.decay_group = DecayGroup(chains)

    def get_decay(self):
        return self.decay_group

    # ------------------------------------------------------------------ parameters
    def _init_params(self):
        """Variables in the creation order of DecayGroup.init_params
        (tf_pwa/amp/core.py:1650-1671) and the constraints of config_loader.py:358-537."""
        vm, g = self.vm, self.decay_group
        for r in g.resonances:
            vm.add_real_var(r.name + "_mass", value=r.mass, trainable=False)
            if r.width is not None:
                vm.add_real_var(r.name + "_width", value=r.width, trainable=False)
        inited = set()
        for c in g.chains:
            vm.add_complex_var(c.params_head() + "_total_0", rng=self.rng)
            for d in c:
                if id(d) in inited:
                    continue
                inited.add(id(d))
                head = d.params_head() + "_g_ls"
                for i in range(len(d.ls)):
                    vm.add_complex_var("{}_{}".format(head, i), rng=self.rng)
                vm.set_fix(head + "_0r", 1.0)
                vm.set_fix(head + "_0i", 0.0)
        cons = (self.config.get("constrains") or {}).get("decay") or {}
        idx = cons.get("fix_chain_idx", 0)
        val = cons.get("fix_chain_val", None)
        if val is None:
            val = self.rng.uniform(0, 2)
        name = g.chains[idx].params_head() + "_total_0"
        vm.set_fix(name + "r", val)
        vm.set_fix(name + "i", 0.0)
        bounds = {}
        for c in g.chains:
            for r in c.inner:
                prop = self.config["particle"].get(r.name, {}) or {}
                fl = str(prop.get("float", "") or "")
                if "m" in fl and r.name + "_mass" not in vm.trainable_vars:
                    vm.set_fix(r.name + "_mass", unfix=True)
                    lo, hi = prop.get("mass_min"), prop.get("mass_max")
                    if lo is not None or hi is not None:
                        bounds[r.name + "_mass"] = (lo, hi)
                if "g" in fl and r.width is not None and r.name + "_width" not in vm.trainable_vars:
                    vm.set_fix(r.name + "_width", unfix=True)
                    lo, hi = prop.get("width_min"), prop.get("width_max")
                    if lo is not None or hi is not None:
                        bounds[r.name + "_width"] = (lo, hi)
        vm.set_bound(bounds)

    def get_params(self, trainable_only=False):
        return self.vm.get_all_dic(trainable_only)

    def set_params(self, params):
        if isinstance(params, str):
            with open(params) as f:
                params = json.load(f)
            params = params.get("value", params)
        self.vm.set_all({k: v for k, v in params.items() if k in self.vm.variables})

    def save_params(self, file_name):
        with open(file_name, "w") as f:
            json.dump({"value": self.vm.get_all_dic()}, f, indent=2)

    # ------------------------------------------------------------------ model and data
    def get_amplitude(self):
        if self.amp is None:
            self.amp = AmplitudeModel(self.decay_group, self.vm,
                                      random_z=self.data_config.get("random_z", True))
        return self.amp

    def get_model(self, w_bkg=None):
        if w_bkg is None:
            w_bkg = self.data_config.get("bg_weight", 1.0)
        return Model(self.get_amplitude(), w_bkg=w_bkg)

    @property
    def dat_order(self):
        return self.data_config.get("dat_order", sorted(self.final_names))

    def data_from_p4(self, p4, weight=None):
        """Upload four-momenta (dict name -> (n, 4), or array in dat_order)."""
        return EventSet(self.get_amplitude(), p4, weight=weight, order=self.dat_order)

    def get_data(self, idx):
        files = self.data_config.get(idx)
        if not files:
            return None
        p4 = load_dat_file(files, len(self.final_names))
        w = None
        wf = self.data_config.get(idx + "_weight")
        if isinstance(wf, (str, list)):
            w = np.concatenate([np.loadtxt(f).reshape(-1) for f in
                                ([wf] if isinstance(wf, str) else wf)])
        return self.data_from_p4(p4, weight=w)

    def get_all_data(self):
        return [self.get_data(k) for k in ("data", "phsp", "bg", "inmc")]

    def get_fcn(self, all_data=None, batch=65000):
        if all_data is None:
            all_data = self.get_all_data()
        data, phsp = all_data[0], all_data[1]
        bg = all_data[2] if len(all_data) > 2 else None
        return FCN(self.get_model(), data, phsp, bg=bg, batch=batch)

    def get_params_error(self, params=None, data=None, phsp=None, bg=None, batch=10000,
                         force_pos=True, fcn=None):
        """Errors from the inverse Hessian (ConfigLoader.get_params_error,
        config_loader.py:814-870 with cal_hesse_error / force_pos_def,
        applications.py:249-311): exact Hessian for couplings. Returns {name: error} and keeps
        the error matrix in ``self.inv_he``."""
        if params is not None:
            self.set_params(getattr(params, "params", params))
        if fcn is None:
            if data is None:
                data, phsp, bg, _ = self.get_all_data()
            fcn = FCN(self.get_model(), data, phsp, bg=bg, batch=batch)
        _, _, h = fcn.nll_grad_hessian()
        self.inv_he = inverse_hessian(h, force_pos=force_pos)
        err = np.sqrt(np.fabs(self.inv_he.diagonal()))
        return dict(zip(self.vm.trainable_vars, err.tolist()))

    def cal_fitfractions(self, mc=None, res=None, gram=None, error_matrix=None):
        """Fit fractions and their errors (tf_pwa/fitfractions.py:178-258) from the MC Gram
        matrix; ``error_matrix`` defaults to the last get_params_error result."""
        from .fitfractions import cal_fitfractions

        if error_matrix is None:
            error_matrix = getattr(self, "inv_he", None)
        return cal_fitfractions(self.get_amplitude(), mc, res=res, gram=gram,
                                error_matrix=error_matrix)

    def fit(self, data=None, phsp=None, bg=None, batch=65000, method="BFGS", maxiter=None,
            print_init_nll=False):
        """ConfigLoader.fit (config_loader.py:744-812). method: a scipy.optimize.minimize
        method, or "native" for the engine's own BFGS (fit_native)."""
        from .fit import fit_native, fit_scipy

        if data is None:
            data, phsp, bg, _ = self.get_all_data()
        fcn = FCN(self.get_model(), data, phsp, bg=bg, batch=batch)
        if method == "native":
            return fit_native(fcn, maxiter=maxiter)
        return fit_scipy(fcn, method=method, maxiter=maxiter)
