"""Pin the FP64 adjoint oracle (oracle/kspace_adjoint.py) with central finite differences.

The functional is J(c) = 1/2 sum (y(c) - d)^2 over the recorded sensor traces of
OracleSimulation (the FP64 restatement of kwave/solvers/kspace_solver.py:691-747).
Cells are drawn away from max(c0), so that c_ref = max(c0) (kspace_solver.py:231),
which the adjoint holds fixed, does not move; two of them are source points, which
carry the source-scale term 2 dt / (D c d_i) (kspace_solver.py:619-624).
"""
import copy

import numpy as np
import pytest


def adjoint_case(shape=(32, 40), nt=90, seed=7, mode="additive", p0=False, line_source=True, alpha=None, bona=None,
                 amplitude=1e4, u_source=None):
    from kwave_amd.data import Vector
    from kwave_amd.kgrid import kWaveGrid
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.utils.signals import tone_burst
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    dx = 1e-4
    kg = kWaveGrid(Vector(list(shape)), Vector([dx] * len(shape)))
    kg.setTime(nt, 0.3 * dx / 1700.0)
    c = 1500.0 + 150.0 * gaussian_filter(rng.random(shape), 2)
    rho = 1000.0 + 100.0 * gaussian_filter(rng.random(shape), 2)
    med = kWaveMedium(sound_speed=c, density=rho)
    if bona is not None:
        med.BonA = bona
    if alpha is not None:  # (alpha_coeff [dB/(MHz^y cm)], alpha_power, alpha_mode)
        med.alpha_coeff, med.alpha_power = alpha[0], alpha[1]
        if alpha[2]:
            med.alpha_mode = alpha[2]
    src = kSource()
    if line_source:
        mask = np.zeros(shape, dtype=bool)
        mask[(8,) + tuple(slice(s // 2 - 3, s // 2 + 3) for s in shape[1:])] = True
        src.p_mask = mask
        src.p = amplitude * tone_burst(1.0 / kg.dt, 2.5e6, 2)
        src.p_mode = mode
    if p0:  # an initial pressure blob beside the line source
        grid = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
        r2 = sum((g - s // 3) ** 2 for g, s in zip(grid, shape))
        src.p0 = 0.2 * amplitude * np.exp(-r2 / 8.0)
    if u_source:  # a velocity source (mode u_source) on a short line, every axis driven
        umask = np.zeros(shape, dtype=bool)
        umask[(12,) + tuple(slice(s // 2 - 2, s // 2 + 2) for s in shape[1:])] = True
        src.u_mask = umask
        usig = amplitude / (1000.0 * 1500.0) * tone_burst(1.0 / kg.dt, 2.5e6, 2)
        for i, name in enumerate(["ux", "uy", "uz"][: len(shape)]):
            setattr(src, name, (1.0 - 0.3 * i) * usig)
        src.u_mode = u_source
    smask = np.zeros(shape, dtype=bool)
    smask[(shape[0] - 9,) + tuple(slice(s // 4, 3 * s // 4, 2) for s in shape[1:])] = True
    sen = kSensor(smask)
    sen.record = ["p"]
    return kg, med, src, sen, tuple(4 for _ in shape), tuple(2.0 for _ in shape)


def _oracle(case, c):
    from oracle.kspace_ref import OracleSimulation

    kg, med, src, sen, pml, alpha = case
    m = copy.deepcopy(med)
    m.sound_speed = c
    return OracleSimulation(kg, m, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha)


def _cells(case, c, n_rand=6, seed=3):
    rng = np.random.default_rng(seed)
    flat = c.ravel()
    ok = np.where(flat < flat.max() - 1.0)[0]
    cells = list(rng.choice(ok, size=n_rand, replace=False))
    for name in ("p_mask", "u_mask"):
        if getattr(case[2], name, None) is not None:
            src_pts = np.where(np.asarray(getattr(case[2], name)).ravel())[0]
            cells += list(src_pts[[0, src_pts.size // 2]])
    return cells


# (alpha_coeff [dB/(MHz^y cm)], y, alpha_mode).  The explicit step is stable for Stokes
# absorption only while tau c^2 k_max^2 dt stays below about 2, which on these grids
# (dx = 0.1 mm, CFL 0.3) means alpha_coeff below about 0.04 dB/(MHz^2 cm).
ABSORBING = {"powerlaw": (40.0, 1.5, None), "powerlaw-no-dispersion": (40.0, 1.1, "no_dispersion"), "stokes": (0.02, 2.0, None),
             "powerlaw-mild": (5.0, 1.5, None)}


# B/A 6 at a source amplitude that makes the nonlinear terms matter: 3e7 Pa puts
# B/A S / rho0 at a few per cent near the source
NONLINEAR = dict(bona=6.0, amplitude=3e7)


@pytest.mark.parametrize("mode,p0,alpha,nonlin", [("additive", False, None, False), ("additive-no-correction", False, None, False),
                                                  ("additive", True, None, False), ("additive", False, "powerlaw", False),
                                                  ("additive", False, "powerlaw-no-dispersion", False), ("additive", True, "stokes", False),
                                                  ("additive", False, None, True), ("additive", True, "powerlaw-mild", True),
                                                  ("additive", False, "stokes", True), ("dirichlet", False, None, False),
                                                  ("dirichlet", True, "powerlaw-mild", True), ("u:additive", False, None, False),
                                                  ("u:additive-no-correction", False, None, False), ("u:dirichlet", False, None, False),
                                                  ("u:additive", True, "powerlaw-mild", True)])
def test_oracle_adjoint_matches_finite_differences(mode, p0, alpha, nonlin):
    """mode "u:<m>": a velocity source in mode <m> beside the additive pressure line source."""
    from oracle.kspace_adjoint import sound_speed_gradient

    umode = mode[2:] if mode.startswith("u:") else None
    case = adjoint_case(mode="additive" if umode else mode, p0=p0, alpha=ABSORBING.get(alpha), u_source=umode,
                        **(NONLINEAR if nonlin else {}))
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    rec0 = _oracle(case, c).run()["p"]
    d = rec0 * 0.7 + 0.05 * np.abs(rec0).max() * np.random.default_rng(5).standard_normal(rec0.shape)
    grad, rec = sound_speed_gradient(_oracle(case, c), rec0 - d)
    assert np.allclose(rec, rec0, rtol=0, atol=0)

    def J(cc):
        return 0.5 * float(np.sum((_oracle(case, cc).run()["p"] - d) ** 2))

    h = 1e-2
    for cell in _cells(case, c):
        cp, cm = c.copy(), c.copy()
        cp.ravel()[cell] += h
        cm.ravel()[cell] -= h
        fd = (J(cp) - J(cm)) / (2 * h)
        g = grad.ravel()[cell]
        assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-12 * np.abs(grad).max()) + 1e-9 * np.abs(grad).max(), (cell, g, fd)


def test_oracle_adjoint_planted_sign_error_is_caught():
    """A dropped source-scale term (a plausible slip) fails the same check at a source point."""
    from oracle import kspace_adjoint

    case = adjoint_case()
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    rec0 = _oracle(case, c).run()["p"]
    d = 0.5 * rec0
    orig = kspace_adjoint._source_points
    try:
        kspace_adjoint._source_points = lambda sim: None
        grad, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
    finally:
        kspace_adjoint._source_points = orig
    cell = int(np.where(np.asarray(case[2].p_mask).ravel())[0][0])
    h = 1e-2
    cp, cm = c.copy(), c.copy()
    cp.ravel()[cell] += h
    cm.ravel()[cell] -= h
    J = lambda cc: 0.5 * float(np.sum((_oracle(case, cc).run()["p"] - d) ** 2))  # noqa: E731
    fd = (J(cp) - J(cm)) / (2 * h)
    assert abs(grad.ravel()[cell] - fd) > 1e-3 * abs(fd)


@pytest.mark.parametrize("mode", ["additive", "dirichlet"])
def test_oracle_adjoint_1d_matches_finite_differences(mode):
    """1-D (64 points): a point pressure source beside p0, power law, B/A and an additive
    velocity source, the same finite-difference bar as above."""
    from oracle.kspace_adjoint import sound_speed_gradient

    case = adjoint_case(shape=(64,), mode=mode, p0=True, alpha=ABSORBING["powerlaw"], u_source="additive", **NONLINEAR)
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    rec0 = _oracle(case, c).run()["p"]
    d = 0.7 * rec0
    grad, _ = sound_speed_gradient(_oracle(case, c), rec0 - d)
    J = lambda cc: 0.5 * float(np.sum((_oracle(case, cc).run()["p"] - d) ** 2))  # noqa: E731
    h = 1e-2
    for cell in _cells(case, c):
        cp, cm = c.copy(), c.copy()
        cp.ravel()[cell] += h
        cm.ravel()[cell] -= h
        fd = (J(cp) - J(cm)) / (2 * h)
        g = grad.ravel()[cell]
        assert abs(g - fd) <= 1e-6 * max(abs(fd), 1e-12 * np.abs(grad).max()) + 1e-9 * np.abs(grad).max(), (cell, g, fd)


@pytest.mark.parametrize("slip", ["speed term dropped", "points not cleared"])
def test_oracle_adjoint_planted_dirichlet_slips_are_caught(slip):
    """A dirichlet pressure source overwrites every rho_i at its points with sig / (D c^2):
    dropping that value's c-derivative, or letting the overwritten densities keep their
    adjoint, fails the finite-difference check at a source point."""
    from oracle import kspace_adjoint

    case = adjoint_case(mode="dirichlet")
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    rec0 = _oracle(case, c).run()["p"]
    d = 0.5 * rec0
    orig = kspace_adjoint._dirichlet_source_term

    def planted(rbar, pts, v, cc):
        rb, dc = orig(rbar, pts, v, cc)
        return (rb, 0.0 * dc) if slip == "speed term dropped" else (rbar, dc)

    try:
        kspace_adjoint._dirichlet_source_term = planted
        grad, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
    finally:
        kspace_adjoint._dirichlet_source_term = orig
    cell = int(np.where(np.asarray(case[2].p_mask).ravel())[0][2])
    h = 1e-2
    cp, cm = c.copy(), c.copy()
    cp.ravel()[cell] += h
    cm.ravel()[cell] -= h
    J = lambda cc: 0.5 * float(np.sum((_oracle(case, cc).run()["p"] - d) ** 2))  # noqa: E731
    fd = (J(cp) - J(cm)) / (2 * h)
    assert abs(grad.ravel()[cell] - fd) > 1e-3 * abs(fd), (slip, grad.ravel()[cell], fd)


@pytest.mark.parametrize("umode,slip", [("additive", "speed term dropped"), ("dirichlet", "points not cleared")])
def test_oracle_adjoint_planted_velocity_source_slips_are_caught(umode, slip):
    """Dropping the additive velocity source's scale derivative, or letting the velocity a
    dirichlet source overwrote keep its adjoint, fails the finite-difference check at a
    velocity-source point."""
    from oracle import kspace_adjoint

    case = adjoint_case(u_source=umode)
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    rec0 = _oracle(case, c).run()["p"]
    d = 0.5 * rec0
    orig = kspace_adjoint._velocity_source_term
    planted = (lambda sim, ub, term, t: (ub, 0.0)) if slip == "speed term dropped" else \
        (lambda sim, ub, term, t: (ub, orig(sim, ub, term, t)[1]))
    try:
        kspace_adjoint._velocity_source_term = planted
        grad, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
    finally:
        kspace_adjoint._velocity_source_term = orig
    cell = int(np.where(np.asarray(case[2].u_mask).ravel())[0][1])
    h = 1e-2
    cp, cm = c.copy(), c.copy()
    cp.ravel()[cell] += h
    cm.ravel()[cell] -= h
    J = lambda cc: 0.5 * float(np.sum((_oracle(case, cc).run()["p"] - d) ** 2))  # noqa: E731
    fd = (J(cp) - J(cm)) / (2 * h)
    assert abs(grad.ravel()[cell] - fd) > 1e-3 * abs(fd), (slip, grad.ravel()[cell], fd)


def test_oracle_adjoint_planted_initial_density_slip_is_caught():
    """Dropping the initial-density term -2 p0 / (D c^3) sum rbar, or halving it, fails the
    finite-difference check at a cell inside the initial pressure blob."""
    from oracle import kspace_adjoint

    case = adjoint_case(p0=True)
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    rec0 = _oracle(case, c).run()["p"]
    d = 0.5 * rec0
    J = lambda cc: 0.5 * float(np.sum((_oracle(case, cc).run()["p"] - d) ** 2))  # noqa: E731
    p0 = np.asarray(case[2].p0)
    cell = int(np.argmax(p0.ravel() * (c.ravel() < c.max() - 1.0)))
    h = 1e-2
    cp, cm = c.copy(), c.copy()
    cp.ravel()[cell] += h
    cm.ravel()[cell] -= h
    fd = (J(cp) - J(cm)) / (2 * h)
    good, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
    assert abs(good.ravel()[cell] - fd) <= 1e-6 * abs(fd)
    orig = kspace_adjoint._initial_density_term
    for planted in (lambda sim, c0, rbar: 0.0 * c0, lambda sim, c0, rbar: 0.5 * orig(sim, c0, rbar)):
        try:
            kspace_adjoint._initial_density_term = planted
            bad, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
        finally:
            kspace_adjoint._initial_density_term = orig
        assert abs(bad.ravel()[cell] - fd) > 1e-3 * abs(fd)


@pytest.mark.parametrize("alpha", ["powerlaw", "stokes"])
def test_oracle_adjoint_planted_absorption_slips_are_caught(alpha):
    """Dropping the c-derivative of tau / eta (or halving it), the absorption path into ubar or
    the dispersion path into rbar fails the finite-difference check at the cell where the
    planted gradient moves most."""
    from oracle import kspace_adjoint

    case = adjoint_case(alpha=ABSORBING[alpha])
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    rec0 = _oracle(case, c).run()["p"]
    d = 0.5 * rec0
    J = lambda cc: 0.5 * float(np.sum((_oracle(case, cc).run()["p"] - d) ** 2))  # noqa: E731
    good, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
    orig_term, orig_adj = kspace_adjoint._absorption_speed_term, kspace_adjoint._absorption_adjoint
    plants = {"speed term dropped": ("_absorption_speed_term", lambda sim, c0, w, A, B: 0.0 * w),
              "speed term halved": ("_absorption_speed_term", lambda sim, c0, w, A, B: 0.5 * orig_term(sim, c0, w, A, B)),
              "absorption path dropped": ("_absorption_adjoint", lambda sim, rho0, w: (0.0, orig_adj(sim, rho0, w)[1]))}
    if alpha == "powerlaw":
        plants["dispersion path dropped"] = ("_absorption_adjoint", lambda sim, rho0, w: (orig_adj(sim, rho0, w)[0], w))
    h = 1e-2
    flat = c.ravel()
    for name, (attr, fn) in plants.items():
        setattr(kspace_adjoint, attr, fn)
        try:
            bad, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
        finally:
            kspace_adjoint._absorption_speed_term, kspace_adjoint._absorption_adjoint = orig_term, orig_adj
        diff = np.abs(bad - good).ravel() * (flat < flat.max() - 1.0)
        cell = int(np.argmax(diff))
        cp, cm = c.copy(), c.copy()
        cp.ravel()[cell] += h
        cm.ravel()[cell] -= h
        fd = (J(cp) - J(cm)) / (2 * h)
        assert abs(good.ravel()[cell] - fd) <= 1e-6 * abs(fd), (name, good.ravel()[cell], fd)
        assert abs(bad.ravel()[cell] - fd) > 1e-3 * abs(fd), (name, bad.ravel()[cell], fd)


def test_oracle_adjoint_planted_nonlinear_slips_are_caught():
    """Dropping the B/A term of the equation of state or the adjoint of S^{t-1} through the
    mass update's nonlinear factor fails the finite-difference check."""
    from oracle import kspace_adjoint

    case = adjoint_case(**NONLINEAR)
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    rec0 = _oracle(case, c).run()["p"]
    d = 0.5 * rec0
    J = lambda cc: 0.5 * float(np.sum((_oracle(case, cc).run()["p"] - d) ** 2))  # noqa: E731
    good, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
    orig_eos, orig_mass = kspace_adjoint._nonlinear_eos_term, kspace_adjoint._nonlinear_mass_term
    plants = {"eos term dropped": ("_nonlinear_eos_term", lambda sim, rho0, w, S: 0.0 * w),
              "mass term dropped": ("_nonlinear_mass_term", lambda sim, rbar, D: 0.0)}
    h = 1e-2
    flat = c.ravel()
    for name, (attr, fn) in plants.items():
        setattr(kspace_adjoint, attr, fn)
        try:
            bad, _ = kspace_adjoint.sound_speed_gradient(_oracle(case, c), rec0 - d)
        finally:
            kspace_adjoint._nonlinear_eos_term, kspace_adjoint._nonlinear_mass_term = orig_eos, orig_mass
        diff = np.abs(bad - good).ravel() * (flat < flat.max() - 1.0)
        cell = int(np.argmax(diff))
        cp, cm = c.copy(), c.copy()
        cp.ravel()[cell] += h
        cm.ravel()[cell] -= h
        fd = (J(cp) - J(cm)) / (2 * h)
        assert abs(good.ravel()[cell] - fd) <= 1e-6 * abs(fd), (name, good.ravel()[cell], fd)
        assert abs(bad.ravel()[cell] - fd) > 1e-3 * abs(fd), (name, bad.ravel()[cell], fd)


def test_adjoint_rejects_prescaled_source_signals():
    """Signals passed already scaled (the binaries' file route) carry no c0 dependence the
    adjoint could differentiate: kwave_amd.adjoint refuses them before any device work."""
    from kwave_amd.adjoint import sound_speed_gradient

    kg, med, src, sen, pml, alpha = adjoint_case()
    with pytest.raises(ValueError, match="source_scaled"):
        sound_speed_gradient(kg, med, src, sen, np.zeros((1, 1)), pml_size=pml, pml_alpha=alpha, source_scaled=True)
