"""Synthetic workloads of the TT benchmark: Dirichlet-Laplacian (+ diagonal potential) TT operators.

The operator Σ_k I⊗…⊗H_k⊗…⊗I with H_k = tridiag(-1,2,-1) + diag(v_k) has minimal TT rank 2; the rank-3
channel form is the non-minimal assembly with separate L and V channels.
"""
import numpy as np


def laplace_cores(d, n, potentials=None, rank3=False):
    """TT cores (r, n, n, r) of Σ_k I⊗…⊗H_k⊗…⊗I with H_k = tridiag(-1,2,-1) (+ diag(v_k)).
    rank3=True builds the non-minimal rank-3 channel form [[I,0,0],[0,I,0],[L,V_k,I]]."""
    L = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    I = np.eye(n)
    cores = []
    for k in range(d):
        V = np.diag(potentials[k]) if potentials is not None else np.zeros((n, n))
        if rank3:
            # channels: 0 = L placed, 1 = V placed, 2 = nothing placed yet
            if d == 1:
                cores.append((L + V).reshape(1, n, n, 1))
            elif k == 0:
                c = np.zeros((1, n, n, 3)); c[0, :, :, 0] = L; c[0, :, :, 1] = V; c[0, :, :, 2] = I
                cores.append(c)
            elif k == d - 1:
                c = np.zeros((3, n, n, 1)); c[0, :, :, 0] = I; c[1, :, :, 0] = I; c[2, :, :, 0] = L + V
                cores.append(c)
            else:
                c = np.zeros((3, n, n, 3))
                c[0, :, :, 0] = I; c[1, :, :, 1] = I; c[2, :, :, 0] = L; c[2, :, :, 1] = V; c[2, :, :, 2] = I
                cores.append(c)
            continue
        H = L + V
        if d == 1:
            cores.append(H.reshape(1, n, n, 1))
        elif k == 0:
            c = np.zeros((1, n, n, 2)); c[0, :, :, 0] = H; c[0, :, :, 1] = I
            cores.append(c)
        elif k == d - 1:
            c = np.zeros((2, n, n, 1)); c[0, :, :, 0] = I; c[1, :, :, 0] = H
            cores.append(c)
        else:
            c = np.zeros((2, n, n, 2)); c[0, :, :, 0] = I; c[1, :, :, 0] = H; c[1, :, :, 1] = I
            cores.append(c)
    return cores


def batch_problems(count, base_seed=0xBAADF00D, d_range=(6, 20), n_choices=(2, 4, 8, 16), r_range=(4, 64), rhs_rank=4):
    """Independent ALS_SPD problems of uneven size: problem p draws d ~ U{d_range}, n from n_choices and the
    solution rank r ~ U{r_range} (then clipped by TTTensor.random's reduce_to_maximal_ranks) with numpy's
    default_rng(base_seed + p); the operator is the rank-3 Laplacian + U(0,1) potential of laplace_cores, and b
    (rank rhs_rank) and x0 (rank r) are TTTensor.random draws after seed(base_seed + p).
    Returns a list of (A, x0, b, (d, n, r))."""
    from . import tt as _tt
    out = []
    for p in range(count):
        rng = np.random.default_rng(base_seed + p)
        d = int(rng.integers(d_range[0], d_range[1] + 1))
        n = int(rng.choice(n_choices))
        r = int(rng.integers(r_range[0], r_range[1] + 1))
        pots = rng.random((d, n))
        A = _tt.TTOperator.from_components(laplace_cores(d, n, list(pots), rank3=True))
        _tt.seed(base_seed + p)
        b = _tt.TTTensor.random([n] * d, [rhs_rank] * (d - 1))
        x = _tt.TTTensor.random([n] * d, [r] * (d - 1))
        out.append((A, x, b, (d, n, r)))
    return out


def main(argv=None):
    """Runs the uneven batch once and prints its wall time: ``python -m xerus_amd.workloads --problems 512
    --sweeps 2 --precision f64 --threads 0`` (threads 0: the lockstep driver). Meant for profilers, which time the
    kernels of exactly this call."""
    import argparse
    import time
    from . import tt as _tt
    from ._lib import XT_F32, XT_F64, lib
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--problems", type=int, default=512)
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--precision", choices=["f64", "f32"], default="f64")
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args(argv)
    probs = batch_problems(a.problems)
    als = _tt.ALS_SPD.copy()
    als.convergenceEpsilon = 0.0
    als.numHalfSweeps = 2
    als.precision = XT_F64 if a.precision == "f64" else XT_F32
    lib().xt_synchronize()
    t0 = time.perf_counter()
    res, st = _tt.als_round_batch(als, [p[0] for p in probs], [p[1] for p in probs], [p[2] for p in probs], a.sweeps, 1e-12,
                                  threads=a.threads)
    dt = time.perf_counter() - t0
    res = np.asarray(res)
    print(f"{a.problems} problems, {a.sweeps} sweeps, {a.precision}, threads {a.threads}: {dt:.3f} s, "
          f"{a.problems * a.sweeps / dt:.1f} sweeps/s, residual median {np.median(res):.2e}, cg {st['cgIterations']}", flush=True)


if __name__ == "__main__":
    main()
