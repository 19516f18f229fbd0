"""Translation of the product's model descriptors (PhysicalModel, priors, camera
configs) into the FP64 oracle's vocabulary, and the oracle's whole-model
evaluations built from np64: simulate, ForwardProbModel.log_prob and its
finite-difference gradient, lstsq_simulate and BackwardProbModel.log_prob.

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg, never by the product path.
"""
import numpy as np

from oracle import np64
from gigalens_amd import distributions as tfd

KIND_NAME = {1: "Normal", 2: "LogNormal", 3: "Uniform", 4: "TruncatedNormal"}


def oracle_prior(prior):
    """[(path, kind_name, args)] in flatten order."""
    return [(p, KIND_NAME[d.kind], d.args()) for p, d in tfd.leaves(prior)]


def oracle_camera(sim_config, camera=None):
    """np64 camera of a SimulatorConfig (the product camera argument is unused: the
    oracle computes its own coordinates and sub-grid PSF)."""
    return np64.Camera(sim_config.delta_pix, sim_config.num_pix, sim_config.supersample, sim_config.kernel,
                       sim_config.transform_pix2angle)


def nested_from_x(prior, x):
    """(bs, d) physical params -> nested params structure of numpy arrays."""
    cols = [x[:, j] for j in range(x.shape[1])]
    return tfd.pack_as(prior, cols)


def oracle_model_lists(phys_model, nested):
    """Convert nested params to the oracle's (kind, params) lists."""
    has_ll = len(phys_model.lens_light) > 0
    groups = (nested[0], nested[1], nested[2]) if has_ll else (nested[0], [], nested[1])
    lenses = []
    for prof, p in zip(phys_model.lenses, groups[0]):
        if prof._name == "EPL":
            lenses.append(("EPL", dict(p, niter=prof.niter)))
        else:
            lenses.append((prof._name, dict(p)))
    def light(prof, p):
        if prof._name == "SHAPELETS":  # sampled amplitudes amp00.. in p
            return (prof._name, dict(p), prof.n_max, prof.interpolate)
        return (prof._name, dict(p))
    ll = [light(prof, p) for prof, p in zip(phys_model.lens_light, groups[1])]
    src = [light(prof, p) for prof, p in zip(phys_model.source_light, groups[2])]
    return lenses, ll, src


def oracle_simulate(cam, phys_model, nested, per_particle_series=True):
    lenses, ll, src = oracle_model_lists(phys_model, nested)
    if per_particle_series:
        lenses2 = []
        for kind, p in lenses:
            if kind == "EPL":
                p = dict(p)
                niter = p.pop("niter")
                nt = np64.epl_series_length_per_particle(p["e1"], p["e2"], niter)
                lenses2.append(("EPL_PP", (p, nt)))
            else:
                lenses2.append((kind, p))
        lenses = lenses2
    X, Y = cam.x, cam.y
    bx, by = X[None].astype(np.float64), Y[None].astype(np.float64)
    for kind, p in lenses:
        if kind == "EPL_PP":
            ax, ay = np64.epl_deriv(X, Y, n_terms=p[1], **p[0])
        elif kind == "EPL":
            ax, ay = np64.epl_deriv(X, Y, **p)
        else:
            ax, ay = np64.MASS[kind](X, Y, **p)
        bx, by = bx - ax, by - ay
    img = 0.0
    for kind, p, *extra in ll:
        img = img + np64.light_eval(kind, X.astype(np.float64), Y.astype(np.float64), p, *extra)
    for kind, p, *extra in src:
        img = img + np64.light_eval(kind, bx, by, p, *extra)
    img = np.where(np.isnan(img), 0.0, img)
    return np64.conv_pool(img, cam.kernel, cam.ss) * cam.conversion_factor


def oracle_log_prob(cam, phys_model, prior, z, observed, bkg, texp, per_particle_series=True):
    """FP64 ForwardProbModel.log_prob (tf/model.py:53-86) with per-particle EPL series
    (or the reference's batch-coupled series length, epl.py:41)."""
    op = oracle_prior(prior)
    x = np.zeros_like(z, dtype=np.float64)
    lpri = 0.0
    for j, (_, kind, args) in enumerate(op):
        x[:, j], fl = np64.bijector_forward(kind, args, z[:, j])
        lpri = lpri + fl + np64.prior_logpdf(kind, args, x[:, j])
    img = oracle_simulate(cam, phys_model, nested_from_x(prior, x), per_particle_series=per_particle_series)
    ll, chi2 = np64.forward_loglike(img, observed, bkg, texp)
    return ll + lpri, chi2, img


def oracle_grad_richardson(cam, phys_model, prior, z, observed, bkg, texp, threads=8):
    """FP64 gradient of oracle_log_prob w.r.t. z (bs, d): central differences at
    h = 1e-3 and 5e-4 relative (max(|z|, 0.1)), Richardson-extrapolated.  The
    2 x 2 x d batched oracle evaluations run on a thread pool (numpy releases the
    GIL on the large array operations)."""
    from concurrent.futures import ThreadPoolExecutor
    z = np.asarray(z, np.float64)

    def ev(job):
        j, hr, sgn = job
        zz = z.copy()
        zz[:, j] += sgn * hr * np.maximum(np.abs(z[:, j]), 1e-1)
        return oracle_log_prob(cam, phys_model, prior, zz, observed, bkg, texp)[0]
    jobs = [(j, hr, s) for j in range(z.shape[1]) for hr in (1e-3, 5e-4) for s in (1, -1)]
    with ThreadPoolExecutor(threads) as ex:
        vals = list(ex.map(ev, jobs))
    g = np.zeros_like(z)
    for j in range(z.shape[1]):
        fp1, fm1, fp2, fm2 = vals[4 * j:4 * j + 4]
        s = np.maximum(np.abs(z[:, j]), 1e-1)
        cd1 = (fp1 - fm1) / (2 * 1e-3 * s)
        cd2 = (fp2 - fm2) / (2 * 5e-4 * s)
        g[:, j] = (4.0 * cd2 - cd1) / 3.0
    return g


def oracle_nested_prior(prior):
    """Nested (kind_name, args) leaves in the structure of the prior (torch_ref format)."""
    def rec(node):
        if isinstance(node, tfd.JointDistributionSequential):
            return [rec(m) for m in node.model]
        if isinstance(node, tfd.JointDistributionNamed):
            return {k: rec(node.model[k]) for k in sorted(node.model)}
        return (KIND_NAME[node.kind], tuple(node.args()))
    return rec(prior)


def oracle_lstsq_lists(phys_model, nested):
    """Nested params of a least-squares model -> oracle (kind, params[, n_max, interp]) lists."""
    has_ll = len(phys_model.lens_light) > 0
    groups = (nested[0], nested[1], nested[2]) if has_ll else (nested[0], [], nested[1])
    lenses = []
    for prof, p in zip(phys_model.lenses, groups[0]):
        lenses.append(("EPL", dict(p, niter=prof.niter)) if prof._name == "EPL" else (prof._name, dict(p)))

    def light(prof, p):
        if prof._name == "SHAPELETS":
            return ("SHAPELETS", dict(p), prof.n_max, prof.interpolate)
        return (prof._name, dict(p))
    ll = [light(prof, p) for prof, p in zip(phys_model.lens_light, groups[1])]
    src = [light(prof, p) for prof, p in zip(phys_model.source_light, groups[2])]
    return lenses, ll, src


def oracle_lstsq(cam, phys_model, nested, obs, err):
    """FP64 lstsq_simulate (tf/simulator.py:99-165): (stacked, coeffs, image), per-particle EPL series."""
    lenses, ll, src = oracle_lstsq_lists(phys_model, nested)
    pp = []
    for kind, p in lenses:
        if kind == "EPL":
            p = dict(p)
            niter = p.pop("niter")
            pp.append(("EPL", dict(p, niter=niter)))
        else:
            pp.append((kind, p))
    # per-particle series length: evaluate each particle separately
    bs = len(np.atleast_1d(next(iter(nested[0][0].values()))))
    st, co, im = [], [], []
    for i in range(bs):
        def pick(d):
            return {k: (v if k == "niter" else np.atleast_1d(np.asarray(v, np.float64))[i:i + 1])
                    for k, v in d.items()}
        L = [(k, pick(p)) for k, p in pp]
        LL = [(it[0], pick(it[1])) + tuple(it[2:]) for it in ll]
        S = [(it[0], pick(it[1])) + tuple(it[2:]) for it in src]
        stacked = np64.lstsq_components(cam, L, LL, S)
        img, coeffs = np64.lstsq_solve(stacked, obs, err)
        st.append(stacked[0]); co.append(coeffs[0]); im.append(img[0])
    return np.stack(st), np.stack(co), np.stack(im)


def oracle_backward_log_prob(cam, phys_model, prior, z, obs, bkg, texp):
    """FP64 BackwardProbModel.log_prob (tf/model.py:134-163) incl. fldj evaluated at x (model.py:160)."""
    op = oracle_prior(prior)
    x = np.zeros_like(z, dtype=np.float64)
    lpri = 0.0
    for j, (_, kind, args) in enumerate(op):
        x[:, j], _ = np64.bijector_forward(kind, args, z[:, j])
        _, fl_at_x = np64.bijector_forward(kind, args, x[:, j])
        lpri = lpri + fl_at_x + np64.prior_logpdf(kind, args, x[:, j])
    err = np64.backward_err_map(obs, bkg, texp)
    _, _, img = oracle_lstsq(cam, phys_model, nested_from_x(prior, x), obs, err)
    ll, chi2 = np64.backward_loglike(img, obs, err)
    return ll + lpri, chi2, img


def z_to_x(prior, z):
    """Unconstrained z (bs, d) -> physical x (bs, d) through the prior's bijectors, FP64."""
    z = np.asarray(z, np.float64)
    x = np.zeros_like(z)
    for j, (_, kind, args) in enumerate(oracle_prior(prior)):
        x[:, j], _ = np64.bijector_forward(kind, args, z[:, j])
    return x
