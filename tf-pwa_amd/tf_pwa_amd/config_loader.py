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

# Keys of a particle entry the engine implements: quantum numbers, mass and width, the default
# particle model's options (amp/core.py:366-418) and the parameter options of
# add_particle_constraints (config_loader.py:407-537). Keys for plots and labels only are
# accepted and ignored; any other key would change the likelihood and raises.
_PARTICLE_SUPPORTED = {"J", "P", "Par", "C", "spins", "mass", "m0", "width", "g0", "model", "bw",
                       "float", "gauss_constr", "params", "bw_l", "running_width", "width_norm"}
_PARAM_TAILS = ("", "_range", "_sigma", "_free", "_constr", "_min", "_max")
_DISPLAY_KEYS = {"display", "latex", "curve_style", "color", "linestyle", "label",
                 "marker"}
# Keys of a decay entry: LS selection (particle.py:207-248, amp/core.py:1212-1233) and the
# decay model; the other HelicityDecay options (amp/core.py:650-760) raise.
_DECAY_SUPPORTED = {"p_break", "c_break", "l_list", "ls_list", "model"}
_DECAY_MODELS = ("default", "gls-bf")
# Constraint blocks of add_constraints (config_loader.py:258-276) the engine implements.
_CONSTRAINT_KEYS = {"decay", "particle", "fix_var", "free_var", "var_range", "var_equal",
                    "gauss_constr"}
# data options that change the likelihood or the angles and are not implemented (with the
# value that is the reference's default, which is accepted)
_DATA_UNSUPPORTED = {"center_mass": False, "r_boost": True, "align_ref": None,
                     "only_left_angle": False, "identical_particles": None,
                     "cp_particles": None, "resolution_size": 1, "extra_var": None,
                     "weight_scale": False, "bg_frac": None, "inject_ratio": None,
                     "using_mix_likelihood": False, "negtive_idx": None,
                     "preprocessor": "default", "amp_model": "default",
                     "data_charge": None, "phsp_charge": None, "bg_charge": None}


def _unsupported(what, key, value=None):
    raise NotImplementedError("{} {!r}{} is not implemented by the engine".format(
        what, key, "" if value is None else " = {!r}".format(value)))


def check_particle_keys(name, prop):
    for k, v in prop.items():
        if k in _PARTICLE_SUPPORTED or k in _DISPLAY_KEYS:
            continue
        base = k
        for pre, full in (("m_", "mass_"), ("g_", "width_")):
            if k.startswith(pre):
                base = full + k[len(pre):]
        if any(base in (v0 + t for t in _PARAM_TAILS) for v0 in ("mass", "width")):
            continue
        _unsupported("particle option of {}:".format(name), k, v)


def check_decay_keys(core, params):
    for k, v in params.items():
        if k in _DISPLAY_KEYS:
            continue
        if k not in _DECAY_SUPPORTED:
            _unsupported("decay option of {}:".format(core), k, v)
        if k == "model" and v not in _DECAY_MODELS:
            _unsupported("decay model of {}:".format(core), k, v)


def particle_params(prop):
    """The parameter options of one particle as add_particle_constraints collects them
    (config_loader.py:418-437): the ``params`` dict, m0/g0 and m_*/g_* renamed to
    mass/width, and keys that start with mass or width."""
    out = dict(prop.get("params") or {})
    for name, v in prop.items():
        for a, b in (("m0", "mass"), ("g0", "width"), ("m_", "mass_"), ("g_", "width_")):
            if name.startswith(a):
                out[b + name[len(a):]] = v
        if name.startswith("mass") or name.startswith("width"):
            out[name] = v
    return out


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
        if "$include" in pconf:
            _unsupported("particle", "$include")
        for block in (top, finals, pconf):
            for name, prop in block.items():
                if not isinstance(prop, (dict, type(None))):
                    _unsupported("particle map (a list of candidates) for", name)
                check_particle_keys(name, prop or {})
                props[name] = {_PARTICLE_KEYS.get(k, k): v for k, v in (prop or {}).items()
                               if k not in _DISPLAY_KEYS}
        self.particle_props = props
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
                check_decay_keys(core, params)
                params = {k: v for k, v in params.items() if k not in _DISPLAY_KEYS}
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
        constrains = self.config.get("constrains") or {}
        for k in constrains:
            if k not in _CONSTRAINT_KEYS:
                _unsupported("constraint block", k)
        self.gauss_constr = {}
        bounds = {}
        # add_decay_constraints (config_loader.py:358-389): chain-0 total and decay_d
        cons = constrains.get("decay") or {}
        for k in cons:
            if k not in ("fix_chain_idx", "fix_chain_val", "decay_d"):
                _unsupported("decay constraint", k)
        idx = cons.get("fix_chain_idx", 0)
        val = cons.get("fix_chain_val", None)
        if val is None:
            val = self.rng.uniform(0, 2)
        name = g.chains[idx].params_head() + "_total_0"
        vm.set_fix(name + "r", val)
        vm.set_fix(name + "i", 0.0)
        if cons.get("decay_d") is not None:
            self._set_decay_d(cons["decay_d"])
        # add_particle_constraints (config_loader.py:407-537)
        for c in g.chains:
            for r in c.inner:
                self._particle_constraints(r, bounds)
        pcons = constrains.get("particle") or {}
        for k in pcons:
            if k != "equal":
                _unsupported("particle constraint", k)
        for attr, groups in (pcons.get("equal") or {}).items():
            if attr not in ("mass", "width"):
                _unsupported("particle equal constraint on", attr)
            for group in groups:
                rs = [r for r in g.resonances if r.name in group]
                vm.set_same(["{}_{}".format(r.name, attr) for r in rs])
        # fix_var, free_var, var_range, var_equal, gauss_constr (config_loader.py:258-318)
        for k, v in (constrains.get("fix_var") or {}).items():
            self._check_var(k)
            vm.set_fix(k, v)
        for k in constrains.get("free_var") or []:
            self._check_var(k)
            vm.set_fix(k, unfix=True)
        for k, v in (constrains.get("var_range") or {}).items():
            self._check_var(k)
            bounds[k] = tuple(v)
        for names in constrains.get("var_equal") or []:
            for k in names:
                self._check_var(k)
            vm.set_same(list(names))
        for k, v in (constrains.get("gauss_constr") or {}).items():
            self._check_var(k)
            self.gauss_constr[k] = tuple(float(x) for x in v)
        vm.set_bound(bounds)

    def _check_var(self, name):
        if name not in self.vm.variables:
            raise KeyError("constraint names unknown variable {!r}".format(name))

    def _set_decay_d(self, decay_d):
        """decay_d of add_decay_constraints (config_loader.py:368-389): one barrier radius for
        every vertex, a list by vertex position in each chain, or a dict by core name. The
        engine keeps one radius per chain (both vertices and the running width), so values
        that differ within a chain are not implemented."""
        g = self.decay_group
        for c in g.chains:
            decs = list(c)
            if isinstance(decay_d, (int, float)):
                vals = [float(decay_d)] * len(decs)
            elif isinstance(decay_d, (list, tuple)):
                vals = [float(x) for x in decay_d[:len(decs)]]
                decs = decs[:len(vals)]
            elif isinstance(decay_d, dict):
                pairs = [(decay_d[d.core.name], d) for _, d in zip(decay_d, decs)
                         if d.core.name in decay_d]
                vals, decs = [float(a) for a, _ in pairs], [d for _, d in pairs]
            else:
                raise ValueError("decay_d should be list or dict")
            for v, d in zip(vals, decs):
                d.d = v
        for c in g.chains:
            ds = {d.d for d in c}
            if len(ds) > 1:
                _unsupported("decay_d with different barrier radii within chain", str(c),
                             sorted(ds))

    def _particle_constraints(self, r, bounds):
        """One resonance's parameter options (set_prefix_constrains, config_loader.py:1129-1200),
        its gauss_constr entry and its float option (config_loader.py:439-500)."""
        vm = self.vm
        prop = self.config["particle"].get(r.name, {}) or {}
        params = particle_params(prop)
        for vname in ("mass", "width"):
            full = "{}_{}".format(r.name, vname)
            if full not in vm.variables:
                continue
            value = params.get(vname)
            if value is not None:
                vm.set(full, value)
            free = params.get(vname + "_free")
            if free:
                vm.set_fix(full, unfix=True)
            elif free is False:
                vm.set_fix(full)
            if params.get(vname + "_range") is not None:
                bounds[full] = tuple(params[vname + "_range"])
            else:
                lo, hi = params.get(vname + "_min"), params.get(vname + "_max")
                if lo is not None or hi is not None:
                    bounds[full] = (lo, hi)
            if params.get(vname + "_constr"):
                sigma = params.get(vname + "_sigma")
                if value is None or sigma is None:
                    raise ValueError("gaussian constraint of {} needs {} and {}_sigma".format(
                        full, vname, vname))
                self.gauss_constr[full] = (float(value), float(sigma))
        for k, sigma in (prop.get("gauss_constr") or {}).items():
            full = "{}_{}".format(r.name, {"m": "mass", "g": "width"}.get(k, k))
            if not sigma:
                raise ValueError("gauss_constr of {} needs a sigma for {}".format(r.name, k))
            self.gauss_constr[full] = (vm.get(full), float(sigma))
        fl = str(prop.get("float", "") or "")
        if "m" in fl and r.name + "_mass" not in vm.trainable_vars:
            vm.set_fix(r.name + "_mass", unfix=True)
            lo, hi = params.get("mass_min"), params.get("mass_max")
            if lo is not None or hi is not None:
                bounds[r.name + "_mass"] = (lo, hi)
        if "g" in fl and r.width is not None and r.name + "_width" not in vm.trainable_vars:
            vm.set_fix(r.name + "_width", unfix=True)
            lo, hi = params.get("width_min"), params.get("width_max")
            if lo is not None or hi is not None:
                bounds[r.name + "_width"] = (lo, hi)

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
        """Model with the background weight of _get_bg_weight (config_loader.py:629-656): the
        number given as bg_weight (0 without one). A weight file is not a number: the loader
        attaches those weights to the background sample itself (get_data), and the model's
        w_bkg only serves background sets that carry no weights."""
        self._check_data_options()
        if w_bkg is None:
            w_bkg = self.data_config.get("bg_weight", 0.0)
            if isinstance(w_bkg, (str, list)):
                w_bkg = 0.0
        model = self.data_config.get("model", "auto")
        if model not in ("auto", "default"):
            _unsupported("data option", "model", model)
        return Model(self.get_amplitude(), w_bkg=float(w_bkg))

    def _check_data_options(self):
        for k, default in _DATA_UNSUPPORTED.items():
            v = self.data_config.get(k, default)
            if v != default:
                _unsupported("data option", k, v)

    @property
    def dat_order(self):
        return self.data_config.get("dat_order", sorted(self.final_names))

    def data_from_p4(self, p4, weight=None):
        """Upload four-momenta (dict name -> (n, 4), or array in dat_order)."""
        return EventSet(self.get_amplitude(), p4, weight=weight, order=self.dat_order)

    def get_data(self, idx):
        """One sample of the data config (SimpleData.get_data / load_data,
        config_loader/data.py:157-170, 232-251): the four-momenta of the listed .dat files
        and the weights of ``<idx>_weight`` (a number for every event or weight files),
        times -1 for background samples (get_weight_sign: names matching "bg*"), which thus
        always carry weights: -bg_weight, or minus the weight file's values."""
        self._check_data_options()
        files = self.data_config.get(idx)
        if not files:
            return None
        p4 = load_dat_file(files, len(self.final_names))
        return self.data_from_p4(p4, weight=self.sample_weight(idx, p4.shape[0]))

    def sample_weight(self, idx, n):
        """The weights get_data attaches to n events of sample ``idx`` (None: no weights)."""
        wf = self.data_config.get(idx + "_weight")
        if isinstance(wf, (str, list)):
            w = np.concatenate([np.loadtxt(f).reshape(-1) for f in
                                ([wf] if isinstance(wf, str) else wf)])[:n]
        elif wf is not None:
            w = np.full(n, float(wf))
        else:
            w = None
        if idx.startswith("bg"):
            w = -(np.ones(n) if w is None else w)
        return w

    def get_all_data(self):
        return [self.get_data(k) for k in ("data", "phsp", "bg", "inmc")]

    def get_fcn(self, all_data=None, batch=65000):
        if all_data is None:
            all_data = self.get_all_data()
        data, phsp = all_data[0], all_data[1]
        bg = all_data[2] if len(all_data) > 2 else None
        return FCN(self.get_model(), data, phsp, bg=bg, batch=batch,
                   gauss_constr=self.gauss_constr)

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
            fcn = FCN(self.get_model(), data, phsp, bg=bg, batch=batch,
                      gauss_constr=self.gauss_constr)
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
        fcn = FCN(self.get_model(), data, phsp, bg=bg, batch=batch,
                  gauss_constr=self.gauss_constr)
        if method == "native":
            return fit_native(fcn, maxiter=maxiter)
        return fit_scipy(fcn, method=method, maxiter=maxiter)
