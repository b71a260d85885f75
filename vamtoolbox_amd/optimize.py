"""Optimiser front end (vamtoolbox/optimize.py:7-217).

``Options`` keeps the reference's keywords and per-method defaults (CAL:
learning_rate 0.01, momentum 0, positivity 0, sigmoid 0.01; OSMO:
inhibition 0; BCLP: response_model default, eps 0.1, weight 1, p 2, q 1,
learning_rate 0.01, g0 None) but, unlike the reference (optimize.py:119-126),
does not mutate class-level default dictionaries across instances.  Extra
BCLP keywords of this package: ``fused`` (default True), ``hip_graph``
(default True), ``gradient`` ("reference" | "exact"), ``log_metrics``.
"""

from __future__ import annotations

from .optimizer import BCLP, CAL, FBP, OSMO, PM


class Options:
    _default_FBP = {"offset": False}
    _default_CAL = {"learning_rate": 0.01, "momentum": 0, "positivity": 0, "sigmoid": 0.01}
    _default_PM = {"rho_1": 1, "rho_2": 1, "p": 1}
    _default_OSMO = {"inhibition": 0}
    _default_BCLP = {"response_model": "default", "eps": 0.1, "weight": 1, "p": 2, "q": 1, "learning_rate": 0.01,
                     "optim_alg": "grad_des", "g0": None}

    def __init__(self, method: str = "CAL", n_iter: int = 50, d_h: float = 0.8, d_l: float = 0.7,
                 filter: str = "ram-lak", units: str = "normalized", blb=0, bub=None, **kwargs):
        self.method = method
        self.n_iter = n_iter
        self.d_h = d_h
        self.d_l = d_l
        self.filter = filter
        self.units = units
        self.blb = blb
        self.bub = bub
        self.__dict__.update(kwargs)
        self.verbose = self.__dict__.get("verbose", False)
        self.save_img_path = self.__dict__.get("save_img_path", None)
        self.bit_depth = self.__dict__.get("bit_depth", None)
        self.exit_param = self.__dict__.get("exit_param", None)

        def pick(defaults):
            d = dict(defaults)
            d.update({k: v for k, v in kwargs.items() if k in d})
            return d

        if method == "FBP":
            self.offset = pick(self._default_FBP)["offset"]
        if method == "CAL":
            d = pick(self._default_CAL)
            self.learning_rate, self.momentum = d["learning_rate"], d["momentum"]
            self.positivity, self.sigmoid = d["positivity"], d["sigmoid"]
        if method == "PM":
            d = pick(self._default_PM)
            self.rho_1, self.rho_2, self.p = d["rho_1"], d["rho_2"], d["p"]
        if method == "OSMO":
            self.inhibition = pick(self._default_OSMO)["inhibition"]
        if method == "BCLP":
            d = pick(self._default_BCLP)
            if isinstance(d["response_model"], str) and d["response_model"] == "default":
                from .response import ResponseModel
                self.response_model = ResponseModel()
            else:
                self.response_model = d["response_model"]
            self.eps, self.weight, self.p, self.q = d["eps"], d["weight"], d["p"], d["q"]
            self.learning_rate, self.optim_alg, self.g0 = d["learning_rate"], d["optim_alg"], d["g0"]
            self.test_alternate_handling = self.__dict__.get("test_alternate_handling", False)

    def __str__(self):
        return str(self.__dict__)


def optimize(target_geo, proj_geo, options: Options, output="packaged"):
    """Run the optimiser selected by ``options.method`` (optimize.py:174-217)."""
    if options.units != "normalized" or proj_geo.absorption_coeff is not None:
        proj_geo.calcAbsorptionMask(target_geo)
    if options.method == "CAL":
        return CAL.minimizeCAL(target_geo, proj_geo, options)
    if options.method == "OSMO":
        return OSMO.minimizeOSMO(target_geo, proj_geo, options)
    if options.method == "BCLP":
        return BCLP.minimizeBCLP(target_geo, proj_geo, options, output)
    if options.method == "FBP":
        return FBP.minimizeFBP(target_geo, proj_geo, options)
    if options.method == "PM":
        return PM.minimizePM(target_geo, proj_geo, options)
    raise ValueError("unknown method %r" % (options.method,))
