"""Run parameters of the 2-D acoustic FWI solver and their text-file format.

The parameter block is the 19 values the reference reads positionally from
its parameter file (geopvi/forward/fwi2d/fdtd2D_modelling.c:1328-1379), in
the same order, written by ``prepare_fwi_parameters`` from an INI ``[FWI]``
section (geopvi/forward/fwi2d/posterior.py:27-68).  ``Lc`` is the configured
half-order; the solver uses ``Lc + 1`` (fdtd2D_modelling.c:70).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import struct

# (INI key, comment line) in file order; the comment lines are the ones the
# reference writes (posterior.py:29-66) so files stay interchangeable.
_INT_FIELDS = (
    ("nx", "nx", "--grid points in x direction (nx)"),
    ("nz", "nz", "--grid points in z direction (nz)"),
    ("pml0", "pml", "--pml points (pml0)"),
    ("Lc", "Lc", "--Finite difference order (Lc)"),
    ("laplace_slover", "laplace_slover",
     "--Method to calculate Laplace operator: 0 for FDM and 1 for PSM, (laplace_slover)"),
    ("ns", "ns", "--Total number of sources (ns)"),
    ("nt", "nt", "--Total time steps (nt)"),
    ("ds", "ds", "--Shot interval in grid points (ds)"),
    ("ns0", "ns0", "--Grid number of the first shot to the left of the model (ns0)"),
    ("depths", "depths", "--Depth of source in grid points (depths)"),
    ("depthr", "depthr", "--Depth of receiver in grid points (depthr)"),
    ("nr", "nr", "--Total number of receivers (nr)"),
    ("dr", "dr", "--Receiver interval in grid points (dr)"),
    ("nr0", "nr0", "--Grid number of the first receiver to the left of the model (nr0)"),
    ("nt_interval", "nt_interval",
     "--Time step interval of saved wavefield during forward (nt_interval)"),
)
_FLOAT_FIELDS = (
    ("dx", "dx", "--Grid spacing in x direction (dx)"),
    ("dz", "dz", "--Grid spacing in z direction (dz)"),
    ("dt", "dt", "--Time step (dt)"),
    ("f0", "f0", "--Dominant frequency (f0)"),
)


def _fd_weights(lc: int):
    """Central second-derivative weights of half-width lc = Lc + 1 (the
    exact rationals of fdtd2D_modelling.c:1272-1326)."""
    table = {4: (-49 / 18, 3 / 2, -3 / 20, 1 / 90),
             5: (-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560),
             6: (-5269 / 1800, 5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150),
             7: (-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925, -1 / 16632)}
    if lc not in table:
        raise ValueError(f"Lc={lc - 1}: supported 3..6")
    return table[lc]


class CParams(ctypes.Structure):
    """C layout of the parameter block (``gpvi_fwi2d_params`` in
    include/geopvi_fwi2d.h): 15 int32 then 4 float32."""

    _fields_ = [(name, ctypes.c_int) for name, _, _ in _INT_FIELDS] + [
        (name, ctypes.c_float) for name, _, _ in _FLOAT_FIELDS
    ]


@dataclasses.dataclass
class FWIParams:
    nx: int
    nz: int
    pml0: int = 10
    Lc: int = 3
    laplace_slover: int = 0
    ns: int = 1
    nt: int = 1000
    ds: int = 1
    ns0: int = 0
    depths: int = 0
    depthr: int = 0
    nr: int = 1
    dr: int = 1
    nr0: int = 0
    nt_interval: int = 2
    dx: float = 10.0
    dz: float = 10.0
    dt: float = 0.001
    f0: float = 10.0

    def __post_init__(self):
        # The solver reads these as float32 (fdtd2D_modelling.c:60-63, read
        # with %f); keep the float32 value so every layer sees the same number.
        for name, _, _ in _FLOAT_FIELDS:
            setattr(self, name, struct.unpack("f", struct.pack("f", float(getattr(self, name))))[0])
        for name, _, _ in _INT_FIELDS:
            setattr(self, name, int(getattr(self, name)))

    # --- derived sizes (fdtd2D_modelling.c:70-90) ---------------------------
    @property
    def lc(self) -> int:
        """Solver half-width (configured Lc + 1)."""
        return self.Lc + 1

    @property
    def pml(self) -> int:
        return self.pml0 + self.lc

    @property
    def ntx(self) -> int:
        return self.nx + 2 * self.pml

    @property
    def ntz(self) -> int:
        return self.nz + 2 * self.pml

    @property
    def ntp(self) -> int:
        return self.ntx * self.ntz

    @property
    def ni(self) -> int:
        return self.nx * self.nz

    @property
    def nsave(self) -> int:
        """Stored wavefields per shot with a full store (fdtd2D_modelling.c:158-160)."""
        return math.ceil(self.nt / self.nt_interval)

    @property
    def record_len(self) -> int:
        """Length of one sample's record vector, laid out [ns][nt][nr]."""
        return self.ns * self.nt * self.nr

    def to_c(self) -> CParams:
        c = CParams()
        for name, _, _ in _INT_FIELDS:
            setattr(c, name, int(getattr(self, name)))
        for name, _, _ in _FLOAT_FIELDS:
            setattr(c, name, float(getattr(self, name)))
        return c

    @classmethod
    def from_c(cls, c: CParams) -> "FWIParams":
        kw = {name: getattr(c, name) for name, _, _ in _INT_FIELDS}
        kw.update({name: float(getattr(c, name)) for name, _, _ in _FLOAT_FIELDS})
        return cls(**kw)

    def replace(self, **kw) -> "FWIParams":
        return dataclasses.replace(self, **kw)

    # --- text formats ---------------------------------------------------------
    def to_text(self) -> str:
        lines = []
        for name, _, comment in _INT_FIELDS:
            lines += [comment, str(int(getattr(self, name)))]
        for name, _, comment in _FLOAT_FIELDS:
            lines += [comment, repr(float(getattr(self, name)))]
        return "\n".join(lines) + "\n"

    def write(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.to_text())

    @classmethod
    def read(cls, path: str) -> "FWIParams":
        """Read the positional file: each value follows one comment line."""
        with open(path) as f:
            text = f.read()
        return cls.parse(text)

    @classmethod
    def parse(cls, text: str) -> "FWIParams":
        # Same reading rule as fdtd2D_modelling.c:1363-1379: skip one line,
        # read one value, skip whitespace (blank lines included).
        pos, n = 0, len(text)
        vals = []
        n_total = len(_INT_FIELDS) + len(_FLOAT_FIELDS)
        for i in range(n_total):
            nl = text.find("\n", pos)
            if nl < 0:
                raise ValueError(f"parameter file ends before value {i + 1} of {n_total}")
            pos = nl + 1
            while pos < n and text[pos].isspace():
                pos += 1
            end = pos
            while end < n and not text[end].isspace():
                end += 1
            tok = text[pos:end]
            if not tok:
                raise ValueError(f"missing value {i + 1} of {n_total}")
            vals.append(tok)
            pos = end
            while pos < n and text[pos].isspace():
                pos += 1
        kw = {}
        for (name, _, _), tok in zip(_INT_FIELDS, vals[: len(_INT_FIELDS)]):
            kw[name] = int(tok)
        for (name, _, _), tok in zip(_FLOAT_FIELDS, vals[len(_INT_FIELDS):]):
            kw[name] = float(tok)
        return cls(**kw)

    @classmethod
    def from_config(cls, config) -> "FWIParams":
        """From a configparser object with an ``[FWI]`` section (cfg.ini layout)."""
        kw = {}
        for name, key, _ in _INT_FIELDS:
            kw[name] = config.getint("FWI", key)
        for name, key, _ in _FLOAT_FIELDS:
            kw[name] = config.getfloat("FWI", key)
        return cls(**kw)

    # --- stability of the time stepping -----------------------------------
    def max_stable_velocity(self) -> float:
        """Largest velocity the explicit scheme is stable for.

        The update u+ = 2u - u- + (v dt)^2 L u (fdtd2D_modelling.c:776) is
        stable when (v dt)^2 * lambda_max(-L) <= 4.  The central
        second-derivative stencil with weights fd_m (fdtd2D_modelling.c:
        1272-1326) has its largest symbol at the Nyquist wavenumber:
        s = -(fd_0 + 2 sum_m fd_m (-1)^m), so lambda_max(-L) =
        s (1/dx^2 + 1/dz^2); the pseudo-spectral Laplacian (laplace_slover=1)
        has s = pi^2.  For half-order 3 at dx = dz this is v dt/dx <= 0.575."""
        if self.laplace_slover == 1:
            s = math.pi ** 2
        else:
            fd = _fd_weights(self.lc)
            s = -(fd[0] + 2.0 * sum(w * (-1) ** m for m, w in enumerate(fd) if m > 0))
        return 2.0 / (self.dt * math.sqrt(s * (1.0 / self.dx ** 2 + 1.0 / self.dz ** 2)))

    # --- geometry (fdtd2D_modelling.c:118-152) -------------------------------
    def source_ix(self):
        return [is_ * self.ds + self.pml + self.ns0 for is_ in range(self.ns)]

    @property
    def source_iz(self) -> int:
        return self.pml + self.depths

    @property
    def receiver_iz(self) -> int:
        return self.pml + self.depthr

    def receiver_ix(self):
        return [self.pml + ir * self.dr + self.nr0 for ir in range(self.nr)]


def prepare_fwi_parameters(paramfile: str, config) -> None:
    """Write the solver's parameter file from an INI ``[FWI]`` section.

    Same layout as geopvi/forward/fwi2d/posterior.py:27-68: one comment line
    before each value; values are copied as the INI strings.
    """
    lines = []
    for _, key, comment in _INT_FIELDS + _FLOAT_FIELDS:
        lines.append(comment + "\n")
        lines.append(config.get("FWI", key) + "\n")
    with open(paramfile, "w") as f:
        f.writelines(lines)
