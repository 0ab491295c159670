"""Benchmark corpus: deterministic generation, sha256 manifests, verified
loading.

Format (kept alike with the reference's corpus so either side reads the
other's files; benchmark_corpus/CORPUS_SCHEMA.md, corpus_loader_v1.py:1-96):

- `<id>.npz`: arrays `dates` <U10 (T,), `ids` <U10 (N,), `names` <U20
  (F,), `factors` f32 (T,N,F), `factor_a` f32 (T,N) = factors[..., 0],
  `returns` f32 (T,N), `price` f64 (T,N), `mask` bool (T,N),
  `forward_returns` f64 (T,N) (h = 5, lag = 1: entry at t+1, exit at
  t+6, the last 6 rows NaN), `h`, `lag` int64 (1,), `schema_version`,
  `generator_version` (0-d strings);
- `<id>.manifest.json`: `hash.data_sha256` = SHA-256 of the npz file
  bytes (upper-case hex), `arrays[*].array_sha256` = SHA-256 of each
  array's C-ordered bytes, shapes, generation parameters and seeds.

The loader is the only read path: it checks the file hash, then every
array hash, and hands out read-only arrays. Leading-prefix subsets give
smaller panels without touching the RNG.

The generator is this project's own model of the benchmark configurations
(seeded, SHA-256-derived per-role streams, written with fixed zip
timestamps so the same call writes the same bytes):

- factors: correlated through the Cholesky factor of an LKJ(eta = 2)
  correlation matrix, N(0,1) innovations, rounded to 3 decimals (about 2 %
  exact ties);
- mask: per-date valid fraction ~ U(0.6, 1.0), listing gaps (start date ~
  U(0, T/2)) and 1 % random holes, with prices carried through masked
  cells so a masked cell is finite (the mask is authoritative);
- returns: one-factor market model, beta ~ U(0.5, 1.5), market ~
  N(0, 0.01), Student-t(4) idiosyncratic noise scaled 0.02; price =
  100 * cumulative product.
"""
from __future__ import annotations

import argparse
import datetime
import hashlib
import io
import json
import pathlib
import zipfile

import numpy as np

SCHEMA_VERSION = "v1"
GENERATOR_VERSION = "1.0.0"
MASTER_SEED = 20261203
H, LAG = 5, 1


class CorpusError(RuntimeError):
    """A corpus file is missing, altered or does not match its manifest."""


def role_rng(master: int, role: str, extra: str = "") -> np.random.Generator:
    """Independent stream per generation role: the SHA-256 digest of
    "master:role:extra" as uint32 entropy (the reference's seed derivation,
    benchmark_corpus/seeds.json `derivation`)."""
    digest = hashlib.sha256(f"{master}:{role}:{extra}".encode("utf-8")).digest()
    return np.random.default_rng(np.frombuffer(digest, dtype=np.uint32).copy())


def _sha256_bytes(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest().upper()


def array_sha256(a: np.ndarray) -> str:
    return _sha256_bytes(np.ascontiguousarray(a).tobytes())


def _lkj_onion(F: int, eta: float, rng: np.random.Generator) -> np.ndarray:
    """Random F x F correlation matrix from LKJ(eta) (onion method)."""
    if F == 1:
        return np.ones((1, 1))
    beta = eta + (F - 2) / 2.0
    r12 = 2.0 * rng.beta(beta, beta) - 1.0
    C = np.array([[1.0, r12], [r12, 1.0]])
    for k in range(2, F):
        beta -= 0.5
        y = rng.beta(k / 2.0, beta)
        u = rng.standard_normal(k)
        u /= np.linalg.norm(u)
        w = np.sqrt(y) * u
        A = np.linalg.cholesky(C)
        z = A @ w
        C = np.block([[C, z[:, None]], [z[None, :], np.ones((1, 1))]])
    return C


def trading_dates(T: int) -> np.ndarray:
    """T weekdays from 2020-01-02."""
    out, d = [], datetime.date(2020, 1, 2)
    while len(out) < T:
        if d.weekday() < 5:
            out.append(d.isoformat())
        d += datetime.timedelta(days=1)
    return np.array(out, dtype="<U10")


def forward_returns(price: np.ndarray, mask: np.ndarray, h: int = H, lag: int = LAG) -> np.ndarray:
    """fwd[t] = price[t+h+lag] / price[t+lag] - 1 where both ends are valid
    (mask and finite); NaN otherwise and in the last h+lag rows."""
    T = price.shape[0]
    fwd = np.full(price.shape, np.nan)
    if T > h + lag:
        entry = price[lag:T - h]
        exit_ = price[lag + h:T]
        ok = mask[lag:T - h] & mask[lag + h:T] & np.isfinite(entry) & np.isfinite(exit_)
        with np.errstate(invalid="ignore", divide="ignore"):
            r = exit_ / entry - 1.0
        fwd[:T - h - lag] = np.where(ok, r, np.nan)
    return fwd


def generate(T: int, N: int, F: int, master: int = MASTER_SEED) -> dict:
    """The arrays of one synthetic corpus of shape (T, N, F)."""
    if T < 1 or N < 1 or F < 1:
        raise ValueError("T, N and F must be >= 1")
    rf = role_rng(master, "factors", f"{T}x{N}x{F}")
    rm = role_rng(master, "mask", f"{T}x{N}")
    rp = role_rng(master, "price", f"{T}x{N}")
    L = np.linalg.cholesky(_lkj_onion(F, 2.0, rf))
    z = rf.standard_normal((T * N, F))
    factors = np.round((z @ L.T).reshape(T, N, F), 3).astype(np.float32)
    frac = rm.uniform(0.6, 1.0, size=(T, 1))
    start = rm.integers(0, max(1, T // 2), size=N)
    mask = (rm.random((T, N)) < frac) & (np.arange(T)[:, None] >= start[None, :])
    mask &= rm.random((T, N)) >= 0.01
    beta = rp.uniform(0.5, 1.5, size=N)
    market = rp.normal(0.0, 0.01, size=(T, 1))
    idio = 0.02 * rp.standard_t(4.0, size=(T, N))
    ret = (beta[None, :] * market + idio).astype(np.float64)
    price = 100.0 * np.cumprod(1.0 + ret, axis=0)
    return {
        "dates": trading_dates(T),
        "ids": np.array([f"SYN{i:05d}" for i in range(N)], dtype="<U10"),
        "names": np.array([f"factor_{k}" for k in range(F)], dtype="<U20"),
        "factors": factors,
        "factor_a": np.ascontiguousarray(factors[:, :, 0]),
        "returns": ret.astype(np.float32),
        "price": price,
        "mask": mask,
        "forward_returns": forward_returns(price, mask),
        "h": np.array([H], dtype=np.int64),
        "lag": np.array([LAG], dtype=np.int64),
        "schema_version": np.array(SCHEMA_VERSION),
        "generator_version": np.array(GENERATOR_VERSION),
    }


def _npz_bytes(arrays: dict) -> bytes:
    """npz of the arrays in insertion order, stored uncompressed with fixed
    member timestamps, so equal arrays give equal bytes."""
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w", compression=zipfile.ZIP_STORED) as zf:
        for k, v in arrays.items():
            one = io.BytesIO()
            np.lib.format.write_array(one, np.asanyarray(v), allow_pickle=False)
            info = zipfile.ZipInfo(k + ".npy", date_time=(1980, 1, 1, 0, 0, 0))
            info.compress_type = zipfile.ZIP_STORED
            zf.writestr(info, one.getvalue())
    return buf.getvalue()


def write(corpus_id: str, arrays: dict, out_dir, params: dict | None = None) -> dict:
    """Write `<corpus_id>.npz` and its manifest into out_dir; returns the
    manifest."""
    out = pathlib.Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    data = _npz_bytes(arrays)
    (out / f"{corpus_id}.npz").write_bytes(data)
    T, N, F = arrays["factors"].shape
    manifest = {
        "corpus_id": corpus_id,
        "family": "synthetic",
        "version": SCHEMA_VERSION,
        "shapes": {"T": int(T), "N": int(N), "F": int(F)},
        "arrays": [{"name": k, "dtype": str(np.asarray(v).dtype), "shape": list(np.shape(v)),
                    "array_sha256": array_sha256(np.asarray(v))} for k, v in arrays.items()],
        "generation_params": dict(params or {}),
        "seeds": {"master_seed": (params or {}).get("master", MASTER_SEED),
                  "derivation": "numpy default_rng(uint32 words of SHA-256(\"master:role:extra\"))",
                  "roles": ["factors", "mask", "price"]},
        "labels": {"h": H, "lag": LAG},
        "hash": {"data_sha256": _sha256_bytes(data), "algorithm": "SHA-256"},
    }
    (out / f"{corpus_id}.manifest.json").write_text(json.dumps(manifest, indent=1) + "\n")
    return manifest


def _frozen(a: np.ndarray) -> np.ndarray:
    a = np.array(a, copy=True, order="C")
    a.flags.writeable = False
    return a


def load(corpus_id: str, data_dir) -> tuple[dict, dict]:
    """(arrays, manifest) of a corpus after checking the file's SHA-256 and
    every array's SHA-256 against the manifest; arrays are read-only.
    Missing files, a hash mismatch or a manifest without hashes raise
    CorpusError (reference: corpus_loader_v1.py:36-79)."""
    d = pathlib.Path(data_dir)
    npz, man = d / f"{corpus_id}.npz", d / f"{corpus_id}.manifest.json"
    if not npz.exists():
        raise CorpusError(f"{npz} not found: generate it (python -m fc_corpus) first")
    if not man.exists():
        raise CorpusError(f"{man} not found")
    manifest = json.loads(man.read_text(encoding="utf-8"))
    want = (manifest.get("hash") or {}).get("data_sha256")
    if not want:
        raise CorpusError(f"{man.name} has no hash.data_sha256")
    got = _sha256_bytes(npz.read_bytes())
    if got != want.upper():
        raise CorpusError(f"{npz.name}: sha256 {got} != manifest {want}")
    with np.load(npz, allow_pickle=False) as z:
        arrays = {k: _frozen(z[k]) for k in z.files}
    for spec in manifest.get("arrays", []):
        name = spec["name"]
        if name not in arrays:
            raise CorpusError(f"{npz.name}: array {name} listed in the manifest is missing")
        if array_sha256(arrays[name]).upper() != spec["array_sha256"].upper():
            raise CorpusError(f"{npz.name}: array {name} does not match its array_sha256")
    return arrays, manifest


def subset_prefix(d: dict, t: int | None = None, n: int | None = None,
                  f: int | None = None) -> dict:
    """Leading (T', N', F') prefix of a loaded corpus, read-only; sizes
    beyond the full shape are a ValueError (reference:
    corpus_loader_v1.py:82-96)."""
    T, N, F = d["factors"].shape
    for v, full, lab in ((t, T, "T"), (n, N, "N"), (f, F, "F")):
        if v is not None and not 1 <= v <= full:
            raise ValueError(f"{lab}'={v} outside [1, {full}]")
    t, n, f = t or T, n or N, f or F
    out = {"dates": d["dates"][:t], "ids": d["ids"][:n], "names": d["names"][:f],
           "factors": d["factors"][:t, :n, :f], "factor_a": d["factor_a"][:t, :n],
           "returns": d["returns"][:t, :n], "price": d["price"][:t, :n],
           "mask": d["mask"][:t, :n], "forward_returns": d["forward_returns"][:t, :n],
           "h": d["h"], "lag": d["lag"]}
    return {k: _frozen(v) for k, v in out.items()}


def main(argv=None):
    ap = argparse.ArgumentParser(description="Generate a synthetic benchmark corpus")
    ap.add_argument("--T", type=int, default=1218)
    ap.add_argument("--N", type=int, default=5000)
    ap.add_argument("--F", type=int, default=12)
    ap.add_argument("--master", type=int, default=MASTER_SEED)
    ap.add_argument("--id", default=None, help="corpus id (default synth_TxNxF_v1)")
    ap.add_argument("--out-dir", default=".")
    a = ap.parse_args(argv)
    cid = a.id or f"synth_{a.T}x{a.N}x{a.F}_v1"
    m = write(cid, generate(a.T, a.N, a.F, a.master), a.out_dir,
              {"T": a.T, "N": a.N, "F": a.F, "master": a.master})
    print(json.dumps({"corpus_id": cid, "data_sha256": m["hash"]["data_sha256"]}))


if __name__ == "__main__":
    main()
