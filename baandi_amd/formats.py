"""On-disk formats of the reference CLI for the Gamma SCF around the path.

* ``density.bin`` + ``density.json`` (schema 1): the real-space density as
  little-endian f64 in C order ((i n1 + j) n2 + k, k fastest) with its
  geometry, charge normalisation and SHA-256 (the files the reference's
  tooling reads, tests/reference/README.md:239-245).
* ``results.json`` (schema 1): the run record with the reference's field
  names and units (crates/dft-cli, read by scripts/qe_common.py:80-190).
* Checkpoint v2 (crates/dft-cli/src/checkpoint_io.rs:2290-2520): the
  ``BAANDICP`` envelope (8-byte magic, u64 LE header and payload lengths,
  compact JSON header, payload, SHA-256 of everything before it) holding an
  ``ExactContinuation`` (one state per rank) or a ``ResultExport`` (energy
  bits and SHA-256 of the delivered density and complex orbitals).
* The state codec inside a continuation (crates/dft-engine/src/refscf/
  checkpoint.rs:97-330): u64 LE; f64 as its IEEE bits; bool as u64 0/1;
  vectors and strings length-prefixed; options tagged by a bool; matrices
  as rows, cols and the column-major values; complex numbers as (re, im).

The Gamma continuation state written here follows the order of the
reference's Gamma driver capture (refscf/driver.rs:539-560: header, density
G^2, completed maps, density, input coefficients, mixer, operator and
orbitals, warm-start entry, ethr, last dr2, per-map records) with this
library's own content in each section; the reference's Rust Debug strings
(atoms, options) are replaced by this library's descriptions, so a file is
a continuation for this library, in the reference's container and codec.
"""
from __future__ import annotations

import hashlib
import json
import math
import struct

import numpy as np

MAGIC = b"BAANDICP"
MAX_FILE = 512 * 1024 * 1024       # checkpoint_io.rs:17
MAX_HEADER = 32 * 1024 * 1024      # checkpoint_io.rs:18
MAX_STATE = 512 * 1024 * 1024      # refscf/checkpoint.rs:9
SCHEMA_VERSION = 2
EV_PER_HA = 27.211386245988        # the unit conversion of the reference tooling (scripts/qe_common.py:9)
SQRT2 = math.sqrt(2.0)


class FormatError(ValueError):
    """A file or state that does not decode (the reference's InvalidData)."""


def _hex64(s):
    return isinstance(s, str) and len(s) == 64 and all(c in "0123456789abcdefABCDEF" for c in s)


# ---------------------------------------------------------------------------
# State codec (refscf/checkpoint.rs:97-330).

class StateWriter:
    def __init__(self):
        self.out = bytearray()

    def _room(self, extra):
        if len(self.out) + extra > MAX_STATE:
            raise FormatError("state quota")

    def u64(self, v):
        v = int(v)
        if not 0 <= v < 1 << 64:
            raise FormatError("integer range")
        self._room(8)
        self.out += struct.pack("<Q", v)

    usize = u64

    def f64(self, v):
        self._room(8)
        self.out += struct.pack("<d", float(v))

    def boolean(self, v):
        self.u64(1 if v else 0)

    def string(self, s):
        b = s.encode("utf-8")
        self.u64(len(b))
        self._room(len(b))
        self.out += b

    def f64s(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype="<f8").reshape(-1))
        self.u64(a.size)
        self._room(8 * a.size)
        self.out += a.tobytes()

    def usizes(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype="<u8").reshape(-1))
        self.u64(a.size)
        self._room(8 * a.size)
        self.out += a.tobytes()

    def fixed_usizes(self, a):
        """[usize; N]: no length prefix."""
        for v in a:
            self.u64(v)

    def fixed_f64s(self, a):
        for v in np.asarray(a, dtype=np.float64).reshape(-1):
            self.f64(v)

    def complexes(self, a):
        """Vec<Complex64>: length, then (re, im) per element."""
        a = np.asarray(a, dtype=np.complex128).reshape(-1)
        self.u64(a.size)
        self._room(16 * a.size)
        self.out += np.ascontiguousarray(a.view(np.float64).astype("<f8")).tobytes()

    def real_matrix(self, columns):
        """RealMatrix rows x cols from a (cols, rows) array whose rows are the
        matrix's columns (column-major values)."""
        c = np.ascontiguousarray(np.asarray(columns, dtype="<f8"))
        cols, rows = c.shape
        self.u64(rows)
        self.u64(cols)
        self._room(8 * c.size)
        self.out += c.tobytes()

    def option(self, value, put):
        self.boolean(value is not None)
        if value is not None:
            put(value)

    def bytes(self):
        return bytes(self.out)


class StateReader:
    def __init__(self, data: bytes):
        self.data = memoryview(data)
        self.cursor = 0

    def remaining(self):
        return len(self.data) - self.cursor

    def _take(self, n):
        if n < 0 or self.cursor + n > len(self.data):
            raise FormatError("truncated state or invalid length")
        out = self.data[self.cursor:self.cursor + n]
        self.cursor += n
        return out

    def u64(self):
        return struct.unpack("<Q", self._take(8))[0]

    usize = u64

    def f64(self):
        return struct.unpack("<d", self._take(8))[0]

    def boolean(self):
        v = self.u64()
        if v not in (0, 1):
            raise FormatError("boolean tag")
        return bool(v)

    def string(self):
        n = self.u64()
        try:
            return bytes(self._take(n)).decode("utf-8")
        except UnicodeDecodeError as e:
            raise FormatError("UTF-8") from e

    def _count(self, width):
        n = self.u64()
        if n > self.remaining() // width:
            raise FormatError("array length quota")
        return n

    def f64s(self):
        n = self._count(8)
        return np.frombuffer(bytes(self._take(8 * n)), dtype="<f8").astype(np.float64)

    def usizes(self):
        n = self._count(8)
        return np.frombuffer(bytes(self._take(8 * n)), dtype="<u8").astype(np.int64)

    def fixed_usizes(self, count):
        return [self.u64() for _ in range(count)]

    def fixed_f64s(self, count):
        return np.array([self.f64() for _ in range(count)])

    def complexes(self):
        n = self._count(16)
        return np.frombuffer(bytes(self._take(16 * n)), dtype="<f8").astype(np.float64).view(np.complex128)

    def real_matrix(self):
        rows, cols = self.u64(), self.u64()
        if rows * cols > self.remaining() // 8:
            raise FormatError("matrix shape quota")
        a = np.frombuffer(bytes(self._take(8 * rows * cols)), dtype="<f8").astype(np.float64)
        return a.reshape(cols, rows)

    def option(self, get):
        return get() if self.boolean() else None

    def finish(self):
        if self.cursor != len(self.data):
            raise FormatError("trailing state bytes")


def _finite(a, what):
    if not np.all(np.isfinite(np.asarray(a).view(np.float64) if np.iscomplexobj(a) else a)):
        raise FormatError(f"nonfinite {what}")


# ---------------------------------------------------------------------------
# The BAANDICP envelope and the v2 header (checkpoint_io.rs:2350-2520).

def envelope(header: dict, payload: bytes) -> bytes:
    head = json.dumps(header, separators=(",", ":"), ensure_ascii=False).encode("utf-8")
    if len(head) > MAX_HEADER or len(head) + len(payload) + 56 > MAX_FILE:
        raise FormatError("v2 container quota")
    out = bytearray(MAGIC)
    out += struct.pack("<QQ", len(head), len(payload))
    out += head
    out += payload
    out += hashlib.sha256(out).digest()
    return bytes(out)


def open_envelope(data: bytes):
    """(header bytes, payload bytes) of a BAANDICP file, after the size,
    length and SHA-256 checks."""
    if len(data) < 56 or len(data) > MAX_FILE or data[:8] != MAGIC:
        raise FormatError("checkpoint magic/size")
    h, b = struct.unpack("<QQ", data[8:24])
    if h > MAX_HEADER or h + b + 56 != len(data):
        raise FormatError("checkpoint lengths")
    if hashlib.sha256(data[:-32]).digest() != data[-32:]:
        raise FormatError("checkpoint checksum")
    return data[24:24 + h], data[24 + h:24 + h + b]


def validate_identity(identity: dict):
    """refscf/checkpoint.rs:27-40."""
    ok = (_hex64(identity.get("input_sha256")) and _hex64(identity.get("binary_sha256"))
          and isinstance(identity.get("upf_sha256"), list) and identity["upf_sha256"]
          and all(_hex64(s) for s in identity["upf_sha256"])
          and isinstance(identity.get("platform"), str) and identity["platform"]
          and int(identity.get("threads", 0)) > 0 and int(identity.get("ranks", 0)) > 0)
    if not ok:
        raise FormatError("invalid input/UPF/binary/execution identity")


def _identity_json(identity):
    return {"input_sha256": identity["input_sha256"], "upf_sha256": list(identity["upf_sha256"]),
            "binary_sha256": identity["binary_sha256"], "platform": identity["platform"],
            "threads": int(identity["threads"]), "ranks": int(identity["ranks"])}


def encode_v2(identity: dict, iteration: int, states=(), result=None) -> bytes:
    """An ExactContinuation (one state per rank, no result) or a ResultExport
    (a result record, no states)."""
    validate_identity(identity)
    states = list(states)
    if (result is None and len(states) != identity["ranks"]) or (result is not None and states):
        raise FormatError("v2 kind/rank count")
    refs, payload = [], bytearray()
    for s in states:
        if not s or len(s) > MAX_STATE:
            raise FormatError("state size quota")
        refs.append({"offset": len(payload), "count": len(s)})
        payload += s
    if result is not None:
        _check_result(result, iteration)
    header = {"schema_version": SCHEMA_VERSION,
              "kind": "ResultExport" if result is not None else "ExactContinuation",
              "identity": _identity_json(identity), "iteration": int(iteration), "states": refs,
              "result": result}
    return envelope(header, bytes(payload))


def _check_result(r, iteration):
    keys = ["energy_ha_bits", "density_sha256", "orbitals_sha256", "converged", "iterations",
            "verification_attempts"]
    if list(r.keys()) != keys:
        raise FormatError("v2 result fields")
    energy = struct.unpack("<d", struct.pack("<Q", int(r["energy_ha_bits"])))[0]
    if not math.isfinite(energy) or r["iterations"] != iteration or not (
            _hex64(r["density_sha256"]) and _hex64(r["orbitals_sha256"])):
        raise FormatError("v2 result energy/hash/iteration")


def decode_v2(data: bytes):
    """dict(kind, identity, iteration, states, result) of a v2 file."""
    head, payload = open_envelope(data)
    try:
        h = json.loads(bytes(head).decode("utf-8"))
    except (UnicodeDecodeError, json.JSONDecodeError) as e:
        raise FormatError("v2 header JSON") from e
    if list(h.keys()) != ["schema_version", "kind", "identity", "iteration", "states", "result"]:
        raise FormatError("v2 header fields")
    if h["schema_version"] != SCHEMA_VERSION:
        raise FormatError("v2 schema version")
    identity = h["identity"]
    if list(identity.keys()) != ["input_sha256", "upf_sha256", "binary_sha256", "platform", "threads", "ranks"]:
        raise FormatError("v2 identity fields")
    validate_identity(identity)
    exact = h["kind"] == "ExactContinuation" and h["result"] is None and len(h["states"]) == identity["ranks"]
    export = h["kind"] == "ResultExport" and h["result"] is not None and not h["states"]
    if not (exact or export):
        raise FormatError("v2 kind/rank count")
    if export:
        _check_result(h["result"], h["iteration"])
    cursor, states = 0, []
    for a in h["states"]:
        if list(a.keys()) != ["offset", "count"] or a["offset"] != cursor or a["count"] == 0 \
                or a["offset"] + a["count"] > len(payload):
            raise FormatError("v2 state offset/count")
        states.append(bytes(payload[cursor:cursor + a["count"]]))
        cursor += a["count"]
    if cursor != len(payload):
        raise FormatError("v2 unreferenced bytes")
    return {"kind": h["kind"], "identity": identity, "iteration": h["iteration"], "states": states,
            "result": h["result"]}


# ---------------------------------------------------------------------------
# Delivered orbitals and the ResultExport record.

def gamma_complex_orbitals(system, orbitals):
    """Complex coefficients (bands, n) in basis order from ORC rows: row 0
    is c(0); row g >= 1 of representative i = ceil(g/2) gives w = v/sqrt2 as
    Re c(+G_i) = Re c(-G_i) (g odd) or Im c(+G_i) = -Im c(-G_i) (g even)
    (refscf/lazy.rs:107-147)."""
    x = np.asarray(orbitals, dtype=np.float64)
    plus = system.get("plus_basis")
    minus = system.get("minus_basis")
    bands, n = x.shape
    full = np.zeros((bands, n), dtype=np.complex128)
    full[:, 0] = x[:, 0]
    w = x[:, 1:] / SQRT2
    reps = np.arange(1, (n + 1) // 2)
    full.real[:, plus[reps]] = w[:, 2 * reps - 2]
    full.real[:, minus[reps]] = w[:, 2 * reps - 2]
    full.imag[:, plus[reps]] = w[:, 2 * reps - 1]
    full.imag[:, minus[reps]] = -w[:, 2 * reps - 1]
    return full


def result_record(energy_ha, density, complex_orbitals, converged, iterations, verification_attempts):
    """The v2 ResultExport record: energy bits and SHA-256 of the density's
    and the orbitals' little-endian IEEE bits in grid C order and
    column-major (band by band, re then im) order (checkpoint_io.rs:2741-2800)."""
    e = float(energy_ha)
    if not math.isfinite(e):
        raise FormatError("v2 nonfinite delivered energy")
    d = hashlib.sha256(np.ascontiguousarray(np.asarray(density, dtype="<f8").reshape(-1)).tobytes())
    o = hashlib.sha256(np.ascontiguousarray(np.asarray(complex_orbitals, dtype=np.complex128)
                                            .view(np.float64).astype("<f8")).tobytes())
    return {"energy_ha_bits": struct.unpack("<Q", struct.pack("<d", e))[0], "density_sha256": d.hexdigest(),
            "orbitals_sha256": o.hexdigest(), "converged": bool(converged), "iterations": int(iterations),
            "verification_attempts": int(verification_attempts)}


# ---------------------------------------------------------------------------
# density.bin / density.json (schema 1).

def write_density(directory, density, lattice_rows, volume, electrons):
    """Write density.bin and density.json into an existing directory; never
    overwrites (the reference tooling's no-clobber rule). Returns the metadata."""
    import os
    rho = np.ascontiguousarray(np.asarray(density, dtype="<f8"))
    if rho.ndim != 3 or not np.all(np.isfinite(rho)):
        raise FormatError("density must be a finite 3-D grid")
    raw = rho.tobytes()
    n = rho.size
    meta = {"dtype": "f64", "endianness": "little", "fft_grid": list(rho.shape),
            "lattice_columns_bohr": np.asarray(lattice_rows, dtype=np.float64).T.tolist(),
            "normalization": {"expected_electrons": float(electrons),
                              "integral_electrons": float(rho.sum() * volume / n)},
            "ordering": "C: (i*n2+j)*n3+k; k fastest", "schema_version": 1,
            "sha256": hashlib.sha256(raw).hexdigest(), "units": "electrons per Bohr^3",
            "volume_bohr3": float(volume)}
    for name, body in (("density.bin", raw), ("density.json", (json.dumps(meta, indent=2) + "\n").encode())):
        with open(os.path.join(directory, name), "xb") as f:
            f.write(body)
    return meta


def read_density(directory):
    """(grid, metadata) of density.bin/density.json, after the schema, shape
    and digest checks."""
    import os
    with open(os.path.join(directory, "density.json")) as f:
        meta = json.load(f)
    if meta.get("schema_version") != 1 or meta.get("dtype") != "f64" or meta.get("endianness") != "little" \
            or meta.get("ordering") != "C: (i*n2+j)*n3+k; k fastest":
        raise FormatError("density.json schema")
    with open(os.path.join(directory, "density.bin"), "rb") as f:
        raw = f.read()
    if hashlib.sha256(raw).hexdigest() != meta["sha256"]:
        raise FormatError("density digest")
    dims = tuple(int(v) for v in meta["fft_grid"])
    if len(raw) != 8 * int(np.prod(dims)):
        raise FormatError("density size")
    return np.frombuffer(raw, dtype="<f8").astype(np.float64).reshape(dims), meta


# ---------------------------------------------------------------------------
# results.json (schema 1).

def _both(ha):
    return {"ha": float(ha), "ry": 2.0 * float(ha)}


def results_record(system, run, options, species="X", pseudo_sha256=None, input_sha256=None, tool=None):
    """The run record of a Gamma SCF (``run`` from ``run_scf``) with the
    reference's field names and units: Ha and Ry energies, eigenvalues in Ha
    and eV, lattice as columns, atoms as species + fractional + Cartesian
    positions; per-map records carry dr2_ry, ethr_ry and h_columns as the
    reference driver reports them (crates/dft-cli/src/telemetry.rs:125-145).
    The last map adds the relative charge error of the output density and
    the largest occupied-band residual norm |Hx - theta x| (Ha).
    ``options``: ecut_ha and the SCF settings to record under "scf"."""
    rows = np.asarray(system.lattice_rows, dtype=np.float64)
    cart = system.get("atoms_cart")
    frac = cart @ np.linalg.inv(rows)
    recip = 2.0 * math.pi * np.linalg.inv(rows)  # columns are the reciprocal vectors
    ecut = float(options.get("ecut_ha", 0.0)) if isinstance(options, dict) else 0.0
    e = run["energy"]
    energies = {k: _both(e[k]) for k in ("kinetic", "local", "nonlocal", "hartree", "xc", "ewald", "total")}
    if "band_sum" in e:
        energies["band_sum"] = _both(e["band_sum"])
    if e.get("entropy_term"):
        energies["entropy_term"] = _both(e["entropy_term"])
    hist = [{"iteration": h["iteration"], "dr2_ry": h["dr2_ry"], "ethr_ry": h["ethr_ry"],
             "davidson_iterations": h["davidson_iterations"], "h_columns": h["h_columns"],
             "mixed_precision": bool(h.get("mixed", False)), "warm_start": bool(h.get("warm_start", False))}
            for h in run["history"]]
    eig = [float(v) for v in run["eigenvalues"]]
    occ = [float(v) for v in run["occupations"]]
    pseudo_sha256 = pseudo_sha256 or hashlib.sha256(repr((system.z_valence, system.rc)).encode()).hexdigest()
    return {
        "atoms": [{"cartesian_bohr": cart[i].tolist(), "fractional": frac[i].tolist(), "species": species}
                  for i in range(len(cart))],
        "bands": len(eig),
        "basis": {"fft_grid": [int(v) for v in system.dims], "fft_grid_overridden": False,
                  "n_density_g": int(system.ndensity), "n_plane_waves": int(system.n)},
        "cell": {"lattice_columns_bohr": rows.T.tolist(), "reciprocal_columns_inverse_bohr": recip.tolist(),
                 "volume_bohr3": float(system.volume)},
        "convergence": {"converged": bool(run.get("converged", True)), "history": hist,
                        "iterations": int(run["iterations"]),
                        "last_iteration": None if not hist else dict(
                            hist[-1], relative_charge_error=float(run["relative_charge_error"]),
                            max_eigen_residual_ha=float(run["max_eigen_residual_ha"]))},
        "cutoff": {"ecutwfc_ha": ecut, "ecutwfc_ry": 2.0 * ecut},
        "eigenvalues": {"ev": [v * EV_PER_HA for v in eig], "ha": eig},
        "electrons": float(system.electrons),
        "energies": energies,
        "input_sha256": input_sha256 or hashlib.sha256(repr(sorted(options.items())).encode()).hexdigest(),
        "occupations": occ,
        "schema_version": 1,
        "scf": {k: v for k, v in options.items() if isinstance(v, (int, float, str, bool)) or v is None},
        "species": [{"functional": "SLA PZ", "n_projectors": int(system.nproj // max(system.natoms, 1)),
                     "name": species, "sha256": pseudo_sha256, "z_valence": float(system.z_valence)}],
        "tool": tool or {"name": "baandi_amd", "version": "0.1.0", "git_commit": "unknown"},
        "units": "Ha, Ry, eV, Bohr, seconds and bytes as keyed; dr2_ry in Ry (half-sphere Hartree metric); "
                 "electrons/occupations in electron counts",
        "xc": "lda-pz",
    }


# ---------------------------------------------------------------------------
# The Gamma SCF continuation state.

def _describe(system):
    rows = np.asarray(system.lattice_rows, dtype=np.float64)
    return (f"atoms_cart={np.round(system.get('atoms_cart'), 12).tolist()} z={system.z_valence!r} "
            f"rc={system.rc!r} n={system.n} nproj={system.nproj}"), rows


def encode_gamma_state(system, state) -> bytes:
    """The continuation state of ``run_scf`` at a map boundary (a dict made
    by ``run_scf(..., stop_after=k)``) in the reference codec."""
    w = StateWriter()
    desc, rows = _describe(system)
    w.string(str(SCHEMA_VERSION))
    w.string("gamma")
    w.usize(0)
    w.usize(1)
    w.fixed_usizes(system.dims)
    w.fixed_f64s(rows.T)
    w.string(desc)
    w.string(state["options"])
    w.string(state["profile"])
    w.string("baandi_amd")
    w.f64s(system.get("density_g2"))
    w.usize(state["completed"])
    w.f64s(state["density"])
    w.complexes(state["x"])
    # Mixer: profile, update count, previous (x, f), the difference history
    # oldest first (broyden.rs:466-510).
    w.string(state["profile"])
    w.usize(state["mixer_updates"])
    w.option(state["mixer_prev"], lambda p: (w.complexes(p[0]), w.complexes(p[1])))
    w.u64(len(state["mixer_pairs"]))
    for dx, df in state["mixer_pairs"]:
        w.complexes(df)
        w.complexes(dx)
    # Operator and orbitals (checkpoint.rs:667-688).
    w.usizes(np.arange(system.n))
    w.complexes(system.get("ionic_density_re") + 1j * system.get("ionic_density_im"))
    w.f64s(system.get("diagonal_static"))
    w.f64(system.electrons)
    w.option(state["epoch"], w.u64)
    w.real_matrix(state["orbitals"])
    # Warm-start entry (kept_hx.rs:237-290): enabled, slabs recomputed from
    # X on restore, no verification capture, then (X, HX, (V, epoch, pairs)).
    w.boolean(state["kept_enabled"])
    w.boolean(True)
    w.boolean(False)
    warm = state["warm"]
    w.option(warm, lambda e: (w.real_matrix(e["x"]), w.real_matrix(e["hx"]), w.f64s(e["potential"]),
                              w.u64(e["epoch"]), w.usize(e["pairs"]), w.string(e["op"])))
    w.f64(state["ethr"])
    w.f64(state["last_dr2"])
    w.u64(len(state["history"]))
    for h in state["history"]:
        w.usize(h["iteration"])
        w.f64(h["dr2_ry"])
        w.f64(h["ethr_ry"])
        w.usize(h["davidson_iterations"])
        w.usize(h["h_columns"])
        w.boolean(h["mixed"])
        w.boolean(h["converged"])
        w.boolean(h["warm_start"])
    return w.bytes()


def decode_gamma_state(system, data: bytes, options_text=None, profile_text=None):
    """The state dict of ``encode_gamma_state``, after checking every saved
    invariant against ``system`` (and the options/profile descriptions when
    given): a mismatch raises FormatError("model/options/layout mismatch")."""
    r = StateReader(data)
    desc, rows = _describe(system)

    def same(got, want):
        ok = np.array_equal(np.asarray(got), np.asarray(want)) if not isinstance(want, str) else got == want
        if not ok:
            raise FormatError("model/options/layout mismatch")

    same(r.string(), str(SCHEMA_VERSION))
    same(r.string(), "gamma")
    same(r.usize(), 0)
    same(r.usize(), 1)
    same(r.fixed_usizes(3), list(system.dims))
    same(r.fixed_f64s(9), rows.T.reshape(-1))
    same(r.string(), desc)
    options = r.string()
    profile = r.string()
    if options_text is not None:
        same(options, options_text)
    if profile_text is not None:
        same(profile, profile_text)
    same(r.string(), "baandi_amd")
    same(r.f64s(), system.get("density_g2"))
    st = {"options": options, "profile": profile, "completed": r.usize()}
    st["density"] = r.f64s()
    st["x"] = r.complexes()
    same(r.string(), profile)
    st["mixer_updates"] = r.usize()
    st["mixer_prev"] = r.option(lambda: (r.complexes(), r.complexes()))
    pairs = []
    for _ in range(r.u64()):
        df = r.complexes()
        dx = r.complexes()
        pairs.append((dx, df))
    st["mixer_pairs"] = pairs
    same(r.usizes(), np.arange(system.n))
    same(r.complexes(), system.get("ionic_density_re") + 1j * system.get("ionic_density_im"))
    same(r.f64s(), system.get("diagonal_static"))
    same(r.f64(), system.electrons)
    st["epoch"] = r.option(r.u64)
    st["orbitals"] = r.real_matrix()
    st["kept_enabled"] = r.boolean()
    same(r.boolean(), True)
    if r.boolean():
        raise FormatError("HX capture during verification")
    st["warm"] = r.option(lambda: {"x": r.real_matrix(), "hx": r.real_matrix(), "potential": r.f64s(),
                                   "epoch": r.u64(), "pairs": r.usize(), "op": r.string()})
    st["ethr"] = r.f64()
    st["last_dr2"] = r.f64()
    hist = []
    for _ in range(r.u64()):
        hist.append({"iteration": r.usize(), "dr2_ry": r.f64(), "ethr_ry": r.f64(),
                     "davidson_iterations": r.usize(), "h_columns": r.usize(), "mixed": r.boolean(),
                     "converged": r.boolean(), "warm_start": r.boolean()})
    st["history"] = hist
    r.finish()
    # Counter and shape checks (refscf/checkpoint.rs:733-750, kept_hx.rs:260-275).
    n = system.n
    if st["orbitals"].shape[1] != n or len(st["density"]) != int(np.prod(system.dims)) \
            or len(st["x"]) != system.ndensity:
        raise FormatError("orbital/density shape")
    if len(hist) != st["completed"] or not (0.0 < st["ethr"] <= 1e-2) or (
            st["completed"] > 0 and not (math.isfinite(st["last_dr2"]) and st["last_dr2"] >= 0.0)):
        raise FormatError("iteration/history/ethr state")
    for a, what in ((st["orbitals"], "orbitals"), (st["density"], "density"), (st["x"], "coefficients")):
        _finite(a, what)
    if st["warm"] is not None:
        e = st["warm"]
        if e["x"].shape != st["orbitals"].shape or e["hx"].shape != e["x"].shape \
                or len(e["potential"]) != int(np.prod(system.dims)) or e["pairs"] > (e["x"].shape[0] + 1) // 2 \
                or e["op"] not in ("fp64", "mixed"):
            raise FormatError("Gamma HX shape")
        for a in (e["x"], e["hx"], e["potential"]):
            _finite(a, "warm-start entry")
    return st


# ---------------------------------------------------------------------------
# Files.

def gamma_identity(system, options_text):
    """The v2 identity of a Gamma run of this library: digests of the cell
    and pseudopotential description, of the options and of the library
    binary; one rank, one host thread."""
    from . import lib as L
    desc, rows = _describe(system)
    cell = repr((rows.tolist(), desc, system.dims))
    with open(L.LIB_PATH, "rb") as f:
        binary = hashlib.sha256(f.read()).hexdigest()
    return {"input_sha256": hashlib.sha256((cell + options_text).encode()).hexdigest(),
            "upf_sha256": [hashlib.sha256(repr((system.z_valence, system.rc)).encode()).hexdigest()],
            "binary_sha256": binary, "platform": "hip-gfx950", "threads": 1, "ranks": 1}


def _publish(path, data):
    """Write a new file; an existing destination is never replaced."""
    with open(path, "xb") as f:
        f.write(data)


def write_checkpoint(path, system, options, state):
    """An ExactContinuation file of ``run_scf(..., stop_after=k)["state"]``."""
    from .scf import describe_options
    options_text, _ = describe_options(options)
    data = encode_v2(gamma_identity(system, options_text), state["completed"], [encode_gamma_state(system, state)])
    _publish(path, data)
    return len(data)


def read_checkpoint(path, system, options):
    """The continuation state of a file written by ``write_checkpoint`` for
    the same system, options and library build."""
    from .scf import describe_options
    options_text, profile_text = describe_options(options)
    with open(path, "rb") as f:
        d = decode_v2(f.read())
    if d["kind"] != "ExactContinuation":
        raise FormatError("v2 ResultExport cannot restart")
    if d["identity"] != gamma_identity(system, options_text):
        raise FormatError("input/UPF/binary/execution identity mismatch")
    st = decode_gamma_state(system, d["states"][0], options_text, profile_text)
    if st["completed"] != d["iteration"]:
        raise FormatError("iteration boundary mismatch")
    return st


def export_result(path, system, options, run):
    """A ResultExport file of a finished ``run_scf``."""
    from .scf import describe_options
    options_text, _ = describe_options(options)
    res = result_record(run["energy"]["total"], run["density"],
                        gamma_complex_orbitals(system, run["orbitals"].detach().cpu().numpy()),
                        run["converged"], run["iterations"], run["verification_attempts"])
    _publish(path, encode_v2(gamma_identity(system, options_text), run["iterations"], result=res))
    return res
