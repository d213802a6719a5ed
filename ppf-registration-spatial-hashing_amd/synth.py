"""Seeded synthetic point clouds for the PPF registration workloads.

Every cloud is a pair of float32 arrays (points [n,3], normals [n,3]) built with
numpy's PCG64 generator, so the same seed gives the same bytes on every machine.

Workloads (sizes follow the benchmark definitions):
  * ``plain_small``  -- ellipsoid (1.0/0.6/0.4) united with a 0.8x0.3x0.3 box,
    2,000 model points; scene = the model under a random rigid motion plus 3,000
    clutter points in a 3x3x3 box; xyz noise 0.002*D, normals tilted by 2 deg.
  * ``dataset_like`` -- four superquadric models (3k-8k points) placed on a
    ground plane in a 50,000-point scene, ~30% occluded, plane = 15% of points.
  * ``large_hash``   -- one 30,000-point bumpy superquadric, 200,000-point scene.
  * ``batched``      -- 16 resident models (1k-20k points), 64 scenes (5k-120k).

Model "diameter" D is the largest bounding-box extent, as the reference's
alignment driver computes it (src/alignment.cpp:246-257), and
d_dist = tau_d * D.  The uniform random rotation is Shoemake's method
(reference include/impl/scene_generation.hpp:33-51).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence, Tuple

import numpy as np

Cloud = Tuple[np.ndarray, np.ndarray]


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _normalize(v: np.ndarray) -> np.ndarray:
    n = np.linalg.norm(v, axis=-1, keepdims=True)
    n[n == 0] = 1.0
    return v / n


def bbox_diameter(points: np.ndarray) -> float:
    """Largest bounding-box extent (reference src/alignment.cpp:246-253)."""
    ext = points.max(axis=0) - points.min(axis=0)
    return float(np.float32(max(ext[0], ext[1], ext[2])))


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform SO(3) sample via Shoemake's quaternion construction."""
    x0, x1, x2 = rng.random(3)
    th1, th2 = 2 * math.pi * x1, 2 * math.pi * x2
    r1, r2 = math.sqrt(1 - x0), math.sqrt(x0)
    w, x, y, z = math.sin(th1) * r1, math.cos(th1) * r1, math.sin(th2) * r2, math.cos(th2) * r2
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def perturb(points: np.ndarray, normals: np.ndarray, sigma: float, tilt_deg: float,
            rng: np.random.Generator) -> Cloud:
    """Gaussian xyz noise and a fixed-angle normal tilt about a random axis."""
    p = points + rng.normal(0.0, sigma, size=points.shape)
    n = _normalize(normals)
    r = _normalize(rng.normal(size=n.shape))
    t = _normalize(r - np.sum(r * n, axis=1, keepdims=True) * n)
    a = math.radians(tilt_deg)
    n2 = _normalize(math.cos(a) * n + math.sin(a) * t)
    return p, n2


def transform(points: np.ndarray, normals: np.ndarray, R: np.ndarray, t: np.ndarray) -> Cloud:
    return points @ R.T + t, normals @ R.T


def pose_matrix(R: np.ndarray, t: np.ndarray) -> np.ndarray:
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return T


def as_cloud(points: np.ndarray, normals: np.ndarray) -> Cloud:
    return (np.ascontiguousarray(points, dtype=np.float32),
            np.ascontiguousarray(normals, dtype=np.float32))


# ----------------------------------------------------------------- surfaces
def _sample_ellipsoid(n: int, axes, rng) -> Cloud:
    a, b, c = axes
    gmax = max(b * c, a * c, a * b)
    pts, nrm = [], []
    have = 0
    while have < n:
        u = _normalize(rng.normal(size=(4 * n, 3)))
        g = np.sqrt((b * c * u[:, 0]) ** 2 + (a * c * u[:, 1]) ** 2 + (a * b * u[:, 2]) ** 2)
        keep = rng.random(len(u)) < g / gmax
        u = u[keep]
        p = u * np.array([a, b, c])
        nn = _normalize(p / np.array([a * a, b * b, c * c]))
        pts.append(p)
        nrm.append(nn)
        have += len(p)
    return np.concatenate(pts)[:n], np.concatenate(nrm)[:n]


def _sample_box(n: int, center, size, rng) -> Cloud:
    cx, cy, cz = center
    sx, sy, sz = size
    faces = []  # (axis, sign, area)
    for axis, (d1, d2) in enumerate([(sy, sz), (sx, sz), (sx, sy)]):
        for sign in (-1.0, 1.0):
            faces.append((axis, sign, d1 * d2))
    areas = np.array([f[2] for f in faces])
    which = rng.choice(len(faces), size=n, p=areas / areas.sum())
    uv = rng.random((n, 2)) - 0.5
    p = np.zeros((n, 3))
    nrm = np.zeros((n, 3))
    half = np.array([sx, sy, sz]) / 2
    for k, (axis, sign, _) in enumerate(faces):
        m = which == k
        others = [i for i in range(3) if i != axis]
        p[m, axis] = sign * half[axis]
        p[m, others[0]] = uv[m, 0] * 2 * half[others[0]]
        p[m, others[1]] = uv[m, 1] * 2 * half[others[1]]
        nrm[m, axis] = sign
    return p + np.array([cx, cy, cz]), nrm


def ellipsoid_box_model(n: int, seed: int) -> Cloud:
    """Union of an ellipsoid (1.0/0.6/0.4) and a box 0.8x0.3x0.3 attached at +x."""
    rng = _rng(seed)
    axes = (1.0, 0.6, 0.4)
    center, size = (0.9, 0.0, 0.0), (0.8, 0.3, 0.3)
    # area-proportional split, then drop samples inside the other solid
    pe, ne = _sample_ellipsoid(4 * n, axes, rng)
    pb, nb = _sample_box(4 * n, center, size, rng)
    in_box = np.all(np.abs(pe - np.array(center)) < np.array(size) / 2, axis=1)
    in_ell = np.sum((pb / np.array(axes)) ** 2, axis=1) < 1.0
    pe, ne = pe[~in_box], ne[~in_box]
    pb, nb = pb[~in_ell], nb[~in_ell]
    area_e = 4 * math.pi * ((axes[0] * axes[1]) ** 1.6 + (axes[0] * axes[2]) ** 1.6 +
                            (axes[1] * axes[2]) ** 1.6) ** (1 / 1.6) / 3 ** (1 / 1.6)
    area_b = 2 * (size[0] * size[1] + size[0] * size[2] + size[1] * size[2])
    fe = len(pe) / (4 * n)
    fb = len(pb) / (4 * n)
    we, wb = area_e * fe, area_b * fb
    ne_n = int(round(n * we / (we + wb)))
    ne_n = min(ne_n, len(pe))
    nb_n = n - ne_n
    p = np.concatenate([pe[:ne_n], pb[:nb_n]])
    nn = np.concatenate([ne[:ne_n], nb[:nb_n]])
    perm = rng.permutation(n)
    return as_cloud(p[perm], nn[perm])


def _sq_fn(t, e):
    return np.sign(t) * np.abs(t) ** e


@dataclasses.dataclass
class Superquadric:
    axes: Tuple[float, float, float]
    e1: float
    e2: float
    bumps: np.ndarray | None = None  # [k, 4]: center xyz, sigma
    bump_amp: float = 0.0

    def surface(self, eta: np.ndarray, om: np.ndarray) -> np.ndarray:
        a1, a2, a3 = self.axes
        ce, se = _sq_fn(np.cos(eta), self.e1), _sq_fn(np.sin(eta), self.e1)
        co, so = _sq_fn(np.cos(om), self.e2), _sq_fn(np.sin(om), self.e2)
        p = np.stack([a1 * ce * co, a2 * ce * so, a3 * se], axis=-1)
        if self.bumps is not None:
            n = self.base_normal(eta, om)
            d = np.zeros(len(p))
            for c in self.bumps:
                d += self.bump_amp * np.exp(-np.sum((p - c[:3]) ** 2, axis=1) / (2 * c[3] ** 2))
            p = p + d[:, None] * n
        return p

    def base_normal(self, eta, om) -> np.ndarray:
        a1, a2, a3 = self.axes
        ce, se = _sq_fn(np.cos(eta), 2 - self.e1), _sq_fn(np.sin(eta), 2 - self.e1)
        co, so = _sq_fn(np.cos(om), 2 - self.e2), _sq_fn(np.sin(om), 2 - self.e2)
        return _normalize(np.stack([ce * co / a1, ce * so / a2, se / a3], axis=-1))

    def sample(self, n: int, rng: np.random.Generator) -> Cloud:
        m = max(20 * n, 20000)
        eta = rng.uniform(-math.pi / 2, math.pi / 2, m)
        om = rng.uniform(-math.pi, math.pi, m)
        h = 1e-5
        p = self.surface(eta, om)
        de = (self.surface(eta + h, om) - self.surface(eta - h, om)) / (2 * h)
        do = (self.surface(eta, om + h) - self.surface(eta, om - h)) / (2 * h)
        cr = np.cross(de, do)
        w = np.linalg.norm(cr, axis=1)
        w = np.where(np.isfinite(w), w, 0.0)
        idx = rng.choice(m, size=n, replace=False, p=w / w.sum())
        if self.bumps is None:
            nrm = self.base_normal(eta[idx], om[idx])
        else:
            nrm = _normalize(cr[idx])
            base = self.base_normal(eta[idx], om[idx])
            flip = np.sum(nrm * base, axis=1) < 0
            nrm[flip] *= -1
        return p[idx], nrm


def random_superquadric(rng: np.random.Generator, bumps: int = 0) -> Superquadric:
    axes = tuple(float(x) for x in rng.uniform(0.5, 1.5, 3))
    e1, e2 = (float(x) for x in rng.uniform(0.3, 1.5, 2))
    sq = Superquadric(axes, e1, e2)
    if bumps:
        d = 2 * max(axes)
        sq.bump_amp = 0.05 * d
        eta = rng.uniform(-math.pi / 2, math.pi / 2, bumps)
        om = rng.uniform(-math.pi, math.pi, bumps)
        centers = sq.surface(eta, om)
        sig = rng.uniform(0.05, 0.15, bumps) * d
        sq.bumps = np.concatenate([centers, sig[:, None]], axis=1)
    return sq


def occlude(points: np.ndarray, normals: np.ndarray, frac: float, rng) -> Cloud:
    """Remove the `frac` of points behind a random cutting plane."""
    d = _normalize(rng.normal(size=3))
    s = points @ d
    cut = np.quantile(s, frac)
    keep = s >= cut
    return points[keep], normals[keep]


# ---------------------------------------------------------------- workloads
@dataclasses.dataclass
class Workload:
    name: str
    models: List[Cloud]
    scenes: List[Cloud]
    d_dists: List[float]           # one per model
    rpdf: int
    vote_count_threshold: float
    cpu_clustering: bool
    truths: List[List[np.ndarray]]  # [scene][model] 4x4 ground truth (identity if absent)
    scene_models: List[List[int]]   # model ids present in each scene
    seed: int
    shapes: list | None = None      # the models' superquadrics, when generated from them


def plain_small(seed: int = 1, n_model: int = 2000, n_clutter: int = 3000) -> Workload:
    rng = _rng(seed)
    mp, mn = ellipsoid_box_model(n_model, seed + 1000)
    D = bbox_diameter(mp)
    R = random_rotation(rng)
    t = rng.uniform(-1, 1, 3)
    sp, sn = transform(mp.astype(np.float64), mn.astype(np.float64), R, t)
    sp, sn = perturb(sp, sn, 0.002 * D, 2.0, rng)
    cp = rng.uniform(-1.5, 1.5, (n_clutter, 3))
    cn = _normalize(rng.normal(size=(n_clutter, 3)))
    P = np.concatenate([sp, cp])
    N = np.concatenate([sn, cn])
    perm = rng.permutation(len(P))
    scene = as_cloud(P[perm], N[perm])
    return Workload("plain_small", [(mp, mn)], [scene], [float(np.float32(0.05 * D))], 5, 0.3,
                    True, [[pose_matrix(R, t)]], [[0]], seed)


def _place_on_plane(models_sq, model_clouds, scene_total, plane_frac, noise_rel, occl, rng,
                    model_ids):
    """Instances of the given models at random non-overlapping poses on z=0."""
    k = len(model_ids)
    n_plane = int(round(plane_frac * scene_total))
    n_inst = scene_total - n_plane
    per = [n_inst // k + (1 if i < n_inst % k else 0) for i in range(k)]
    pts, nrm, truths = [], [], []
    spacing = 0.0
    for mid in model_ids:
        spacing = max(spacing, 1.2 * bbox_diameter(model_clouds[mid][0]) + 0.2)
    cols = int(math.ceil(math.sqrt(k)))
    for i, mid in enumerate(model_ids):
        sq = models_sq[mid]
        D = bbox_diameter(model_clouds[mid][0])
        want = int(math.ceil(per[i] / (1.0 - occl)))
        p, n = sq.sample(want, rng)
        p, n = occlude(p, n, occl, rng)
        p, n = p[:per[i]], n[:per[i]]
        R = random_rotation(rng)
        rp = p @ R.T
        gx, gy = i % cols, i // cols
        t = np.array([gx * spacing + rng.uniform(-0.1, 0.1) * spacing,
                      gy * spacing + rng.uniform(-0.1, 0.1) * spacing,
                      -rp[:, 2].min()])
        p2, n2 = transform(p, n, R, t)
        p2, n2 = perturb(p2, n2, noise_rel * D, 2.0, rng)
        pts.append(p2)
        nrm.append(n2)
        truths.append(pose_matrix(R, t))
    lo = np.min(np.concatenate(pts), axis=0) - 0.25 * spacing
    hi = np.max(np.concatenate(pts), axis=0) + 0.25 * spacing
    gp = np.stack([rng.uniform(lo[0], hi[0], n_plane), rng.uniform(lo[1], hi[1], n_plane),
                   np.zeros(n_plane)], axis=1)
    gn = np.tile(np.array([0.0, 0.0, 1.0]), (n_plane, 1))
    gp, gn = perturb(gp, gn, noise_rel * spacing * 0.5, 2.0, rng)
    P = np.concatenate(pts + [gp])
    N = np.concatenate(nrm + [gn])
    perm = rng.permutation(len(P))
    return as_cloud(P[perm], N[perm]), truths


def dataset_like(seed: int = 2, scene_points: int = 50000,
                 model_sizes: Sequence[int] | None = None) -> Workload:
    rng = _rng(seed)
    if model_sizes is None:
        model_sizes = [int(x) for x in rng.integers(3000, 8001, 4)]
    sqs = [random_superquadric(rng) for _ in model_sizes]
    models = [as_cloud(*sq.sample(n, rng)) for sq, n in zip(sqs, model_sizes)]
    taus = rng.uniform(0.03, 0.05, len(models))
    d_dists = [float(np.float32(t * bbox_diameter(m[0]))) for t, m in zip(taus, models)]
    ids = list(range(len(models)))
    scene, truths = _place_on_plane(sqs, models, scene_points, 0.15, 0.003, 0.3, rng, ids)
    return Workload("dataset_like", models, [scene], d_dists, 5, 0.3, False, [truths], [ids],
                    seed)


def large_hash(seed: int = 3, n_model: int = 30000, scene_points: int = 200000) -> Workload:
    rng = _rng(seed)
    sq = random_superquadric(rng, bumps=8)
    model = as_cloud(*sq.sample(n_model, rng))
    D = bbox_diameter(model[0])
    n_clutter = int(0.4 * scene_points)
    inst = scene_points - n_clutter
    per = [inst // 3 + (1 if i < inst % 3 else 0) for i in range(3)]
    pts, nrm, truths = [], [], []
    for i in range(3):
        p, n = sq.sample(per[i], rng)
        R = random_rotation(rng)
        t = np.array([i * 1.5 * D, 0.0, 0.0]) + rng.uniform(-0.1, 0.1, 3) * D
        p2, n2 = transform(p, n, R, t)
        p2, n2 = perturb(p2, n2, 0.002 * D, 2.0, rng)
        pts.append(p2)
        nrm.append(n2)
        truths.append(pose_matrix(R, t))
    allp = np.concatenate(pts)
    lo, hi = allp.min(axis=0) - 0.3 * D, allp.max(axis=0) + 0.3 * D
    cp = rng.uniform(lo, hi, (n_clutter, 3))
    cn = _normalize(rng.normal(size=(n_clutter, 3)))
    P = np.concatenate(pts + [cp])
    N = np.concatenate(nrm + [cn])
    perm = rng.permutation(len(P))
    scene = as_cloud(P[perm], N[perm])
    return Workload("large_hash", [model], [scene], [float(np.float32(0.05 * D))], 5, 0.3, False,
                    [[truths[0]]], [[0]], seed)


def _log_uniform_int(rng, lo, hi, n):
    return [int(round(math.exp(x))) for x in rng.uniform(math.log(lo), math.log(hi), n)]


def batched(seed: int = 4, n_models: int = 16, n_scenes: int = 64,
            model_range=(1000, 20000), scene_range=(5000, 120000),
            models_per_scene=(1, 4)) -> Workload:
    rng = _rng(seed)
    sizes = _log_uniform_int(rng, *model_range, n_models)
    sqs = [random_superquadric(rng) for _ in sizes]
    models = [as_cloud(*sq.sample(n, rng)) for sq, n in zip(sqs, sizes)]
    taus = rng.uniform(0.03, 0.05, n_models)
    d_dists = [float(np.float32(t * bbox_diameter(m[0]))) for t, m in zip(taus, models)]
    scenes, truths, scene_models = [], [], []
    for s in range(n_scenes):
        n_pts = _log_uniform_int(rng, *scene_range, 1)[0]
        k = int(rng.integers(models_per_scene[0], models_per_scene[1] + 1))
        ids = [int(x) for x in rng.choice(n_models, size=k, replace=False)]
        sc, tr = _place_on_plane(sqs, models, n_pts, 0.15, 0.003, 0.3, rng, ids)
        scenes.append(sc)
        truths.append(tr)
        scene_models.append(ids)
    return Workload("batched", models, scenes, d_dists, 5, 0.3, False, truths, scene_models, seed,
                    shapes=sqs)


def served_requests(w: Workload, seed: int = 5, n_requests: int = 32, scene_range=(5000, 100000),
                    targets=(1, 8), keep=None):
    """Requests of the served configuration over the resident models of `w`
    (a `batched` workload): scenes of log-uniform size in scene_range, each
    built like the dataset-like scene (the target models on a ground plane,
    30% occluded, 15% plane points, noise 0.003 D) with 1-8 target model ids.
    keep(n_scene, model_id) -> bool, when given, drops the targets it rejects
    (a request left without targets is drawn again).  Returns [(scene cloud,
    target ids)]."""
    rng = _rng(seed)
    out = []
    while len(out) < n_requests:
        n_pts = _log_uniform_int(rng, *scene_range, 1)[0]
        k = int(rng.integers(targets[0], targets[1] + 1))
        ids = [int(x) for x in rng.choice(len(w.models), size=k, replace=False)]
        if keep is not None:
            ids = [i for i in ids if keep(n_pts, i)]
            if not ids:
                continue
        sc, _ = _place_on_plane(w.shapes, w.models, n_pts, 0.15, 0.003, 0.3, rng, ids)
        out.append((sc, ids))
    return out


WORKLOADS = {
    "plain_small": plain_small,
    "dataset_like": dataset_like,
    "large_hash": large_hash,
    "batched": batched,
}
