"""The drop-in flight-modes park: `VecModesPark`, the reference's batched park API (src/exact_ballpark/vec_modes.py:193-
960 and the setters and OnTick of src/exact_ballpark/vec_orient.py:293-879), on state laid out for the gfx950 kernel.

The park's per-ball state lives in three device blocks, one per element type, one block row per field (structure of
arrays; exact-ballpark_amd/csrc/xb_layout.h). Every attribute the reference's park exposes (`pos` [n, 3], `mass` [n],
`mode` [n], `free` [n] bool, ...) is a view of its block rows with the reference's shape and dtype; assigning one copies
into the block, so the blocks never move and a captured HIP graph keeps its buffers without a re-sync. `evolve()` is
one launch of the HIP kernel (libxb_hip.so, xb_modes_tick) on torch's current stream, no host synchronisation; on a
CPU park it is the host build of the same kernel source (libxb_cpu.so). The commands and setters are the reference's:
elementwise selections between ticks, each a correctly rounded operation in the reference's order.

Worlds: `world` groups the rows; a FOLLOW or ORBIT ball's `follow_ptr` is its target's row, in its world. When every
world's rows are contiguous and no world has more rows than a workgroup (`block`), the park runs the tiled launch
(whole worlds per workgroup, their tick-start records in LDS); otherwise the general launch.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _native as N

F64, I64, U8 = torch.float64, torch.int64, torch.uint8
GOTO, FOLLOW, STOP, WARP, ORBIT, MISSILE, MUSHROOM = 0, 1, 2, 3, 4, 5, 6
BOID, TROLL, MINIBALL, FIELD, RIGID, FORMATION = 7, 8, 9, 10, 11, 12
INERT_MODES = (BOID, MINIBALL, FIELD, RIGID)
LOADABLE = (STOP,) + INERT_MODES
MAX_CATCH_UP_TICKS = 5                      # OnTick's resynchronisation (full.py:92)
FORMATION_MAX_SLOTS = 16                    # a leader's reservations, a std::bitset<16> (Ball.h)

# (name, block, first block row, columns; 0 for a [n] field)
FIELDS = (
    ("pos", "f", 0, 3), ("vel", "f", 3, 3), ("old_pos", "f", 6, 3), ("old_vel", "f", 9, 3), ("goto", "f", 12, 3),
    ("last_g", "f", 15, 3), ("mass", "f", 18, 0), ("time_factor", "f", 19, 0), ("last_collision", "f", 20, 0),
    ("radius", "f", 21, 0), ("max_vel", "f", 22, 0), ("agility", "f", 23, 0), ("speed_fraction", "f", 24, 0),
    ("follow_range", "f", 25, 0), ("max_ang_vel", "f", 26, 0), ("rot_agility", "f", 27, 0), ("rot_roll", "f", 28, 0),
    ("rot", "f", 29, 4), ("old_rot", "f", 33, 4), ("ang_vel", "f", 37, 3), ("old_ang_vel", "f", 40, 3),
    ("torque", "f", 43, 3),
    ("ids", "q", 0, 0), ("mode", "q", 1, 0), ("effect_stamp", "q", 2, 0), ("follow_id", "q", 3, 0),
    ("owner_id", "q", 4, 0), ("follow_ptr", "q", 5, 0), ("clock", "q", 6, 0), ("world", "q", 7, 0),
    ("cloaked", "q", 8, 0), ("bubble", "q", 9, 0),
    ("free", "b", 0, 0), ("moribund", "b", 1, 0), ("massive", "b", 2, 0), ("interactive", "b", 3, 0),
    ("sensor", "b", 4, 0), ("junk", "b", 5, 0), ("formation_id", "c", 6, 0),
)
FIELD_NAMES = tuple(f[0] for f in FIELDS)
# with the gradient collision method: mLastC, the filing (the level and key of the ball's box, its side on each axis),
# the tick's collision list's length
COLL_FIELDS = (("last_c", "cf", 0, 3), ("level", "cq", 0, 0), ("key", "cq", 1, 3), ("side", "cq", 4, 3),
               ("hits_n", "cq", 7, 0))
MAX_COLLIDE_WORLD = 64                      # rows a world with collisions (one wavefront's contact mask)
LOCAL_BALLS = -1073741824                   # DSTLOCALBALLS: a client's local ids count down from it
EMPTY_ID = -(2 ** 62)                       # an empty row's id (a row kept for a ball added mid-run: a mushroom)
MUSHROOM_MASS = float(np.float32(1.0e20))   # AddMushroom's mass, the float literal 1.0e20f (Ballpark.cpp:3742)
MAX_WORLD_STATICS = 64                      # static collidables a world (one response mask)
# with the display orientation kept (Ball::CalculateYawPitchRoll, without dynamical orientation)
YPR_FIELDS = (("yaw", "y", 0, 0), ("pitch", "y", 1, 0), ("roll", "y", 2, 0), ("old_yaw", "y", 3, 0),
              ("old_pitch", "y", 4, 0), ("old_roll", "y", 5, 0), ("roll_speed", "y", 6, 0),
              ("old_roll_speed", "y", 7, 0), ("yaw_delta", "y", 8, 0))
YPR_PATHS = ("host", "cuda")                # the reference's names for where its orientation park evaluates them


# the command interpreter's ops (exact-ballpark_amd/csrc/xb_commands.h, XB_CMD_*) and its record (include/xb.h's
# xb_command: 29 words of 8 bytes; word 28 the gathered state bits, XB_ACCEPTED among them)
CMD = {"goto": 1, "goto_direction": 2, "stop": 3, "follow": 4, "orbit": 5, "warp": 6, "warp_in": 7,
       "stop_release": 8, "release": 9, "set_mass": 10, "set_agility": 11, "set_max_vel": 12,
       "set_speed_fraction": 13, "set_radius": 14, "set_velocity": 15, "set_max_ang_speed": 16,
       "set_rot_agility": 17, "set_ang_vel": 18, "set_rotation": 19, "set_rotation_ypr": 20, "set_free": 21,
       "set_rigid": 22, "set_boid": 23, "set_interactive": 24, "set_troll": 25, "missile_follow": 26,
       "launch_missile": 27}
CMD_WORDS = 29
CMD_STATE_WORD = 28
CMD_ACCEPTED = 32
REFUSALS = {1: "a mushroom as a follow, orbit, missile or formation target is not supported",
            2: "a launcher at rest aims by its heading: without dynamical orientation, that is the display "
               "orientation (make the park with display_ypr=True)"}


def _refusal(status: int) -> None:
    """Raises the gather step's refusal, if any (its status bits)."""
    for bit, msg in REFUSALS.items():
        if status & bit:
            raise ValueError(msg)


def f32(x: torch.Tensor) -> torch.Tensor:
    """A double rounded to the nearest float and back (a C++ float member; out of range an infinity)."""
    return x.to(torch.float32).to(F64)


def _dot(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _sqrt(x: torch.Tensor) -> torch.Tensor:
    """IEEE square root, correctly rounded, by the library's own kernel (or its host build): torch's CPU float64 sqrt
    is not correctly rounded in every case."""
    return N.maths("sqrt", x)


def rotation_ypr(yaw: torch.Tensor, pitch: torch.Tensor, roll: torch.Tensor) -> torch.Tensor:
    """RotationQuaternion(yaw, pitch, roll) of the floats, [n, 4] (x, y, z, w) floats held in doubles: DirectXMath's
    XMQuaternionRotationRollPitchYaw, its scalar path (the reference's q_rotation_ypr, src/exact_ballpark/vec_modes.py:
    184-190, full.py:268-280): each half angle, glibc's cosf and sinf of it, the four sums of float products left to
    right. Each float operation as a double operation rounded to float (the same for * + - of floats)."""
    yaw, pitch, roll = (f32(x.to(F64)) for x in (yaw, pitch, roll))
    hp, hy, hr = f32(pitch * 0.5), f32(yaw * 0.5), f32(roll * 0.5)
    cp, sp = N.maths("cosf", hp), N.maths("sinf", hp)
    cy, sy = N.maths("cosf", hy), N.maths("sinf", hy)
    cr, sr = N.maths("cosf", hr), N.maths("sinf", hr)
    x = f32(f32(f32(cr * sp) * cy) + f32(f32(sr * cp) * sy))
    y = f32(f32(f32(cr * cp) * sy) - f32(f32(sr * sp) * cy))
    z = f32(f32(f32(sr * cp) * cy) - f32(f32(cr * sp) * sy))
    w = f32(f32(f32(cr * cp) * cy) + f32(f32(sr * sp) * sy))
    return torch.stack((x, y, z, w), dim=1)


def rotated_vector(q: torch.Tensor, vec) -> torch.Tensor:
    """Ball::GetRotatedVector's TransformCoord(vec, RotationQuaternion matrix of q) for [n, 4] float quaternions held in
    doubles (full.py:283-324): the matrix's entries and the row vector's products and sums as float operations, divided
    by the homogeneous norm (1 here)."""
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    m00 = f32(1.0 - f32(2.0 * f32(f32(y * y) + f32(z * z))))
    m01 = f32(2.0 * f32(f32(x * y) + f32(z * w)))
    m02 = f32(2.0 * f32(f32(x * z) - f32(y * w)))
    m10 = f32(2.0 * f32(f32(x * y) - f32(z * w)))
    m11 = f32(1.0 - f32(2.0 * f32(f32(x * x) + f32(z * z))))
    m12 = f32(2.0 * f32(f32(y * z) + f32(x * w)))
    m20 = f32(2.0 * f32(f32(x * z) + f32(y * w)))
    m21 = f32(2.0 * f32(f32(y * z) - f32(x * w)))
    m22 = f32(1.0 - f32(2.0 * f32(f32(x * x) + f32(y * y))))
    vx, vy, vz = (float(c) for c in vec)
    zero = torch.zeros_like(x)
    norm = f32(f32(f32(f32(zero * vx) + f32(zero * vy)) + f32(zero * vz)) + 1.0)
    cols = []
    for a, b, c in ((m00, m10, m20), (m01, m11, m21), (m02, m12, m22)):
        cols.append(f32(f32(f32(f32(f32(a * vx) + f32(b * vy)) + f32(c * vz)) + 0.0) / norm))
    out = torch.stack(cols, dim=1)
    return torch.where((norm != 0.0)[:, None], out, 0.0)


def _field_property(name: str, blk: str, row: int, cols: int):
    def get(self):
        if blk == "f":
            base = self._f
        elif blk == "q":
            base = self._q
        elif blk == "cf" or blk == "cq":
            base = self._cf if blk == "cf" else self._cq
            if base is None:
                raise AttributeError(f"{name}: the park was made without collisions")
        elif blk == "y":
            base = self._y
            if base is None:
                raise AttributeError(f"{name}: the park was made without display_ypr")
        elif blk == "c":
            base = self._b.view(torch.int8)
        else:
            base = self._b.view(torch.bool)
        if cols:
            return base[row:row + cols, :self.n].t()
        return base[row, :self.n]

    def set_(self, value):
        dst = get(self)
        if not torch.is_tensor(value):
            value = torch.as_tensor(value)
        dst.copy_(value.to(device=dst.device).expand_as(dst) if value.dim() < dst.dim() else value)

    return property(get, set_, doc=f"[n{', ' + str(cols) if cols else ''}] view of the state block ({name})")


class _Fields(type):
    def __new__(mcs, name, bases, ns):
        for fname, blk, row, cols in FIELDS + COLL_FIELDS + YPR_FIELDS:
            ns.setdefault(fname, _field_property(fname, blk, row, cols))
        return super().__new__(mcs, name, bases, ns)


class VecModesPark(metaclass=_Fields):
    """n balls, one per row, in independent worlds, with destiny's flight modes; `evolve()` is one kernel launch a
    tick. The reference's constructor (src/exact_ballpark/vec_modes.py:199-217).

    `ypr` takes the reference's values ("host", "cuda": where its orientation park evaluates atan2, asin and sin); its
    modes park keeps no yaw, pitch and roll, and neither does this one unless `display_ypr`: then FullBallpark's
    display orientation without dynamical orientation (Ball::CalculateYawPitchRoll, src/exact_ballpark/full.py:
    2043-2089) is kept in a block of its own, snapped when a free ball is loaded and when a velocity is set, evolved
    at every live row's commit inside the tick's kernel, with glibc's functions on the device either way."""

    def __init__(self, n: int, friction: float = 1.0e6, tick_interval_ms: int = 1000, device="cuda",
                 dynamical_orientation: bool = False, use_new_orbit: bool = False, block: int = 256,
                 ypr: str = "host", collisions: bool = False, display_ypr: bool = False, master: bool = False):
        if ypr not in YPR_PATHS:
            raise ValueError(f"ypr must be one of {YPR_PATHS}, not {ypr!r}")
        if display_ypr and dynamical_orientation:
            raise ValueError("display_ypr is the orientation without dynamical orientation")
        self.ypr = ypr
        self.master = bool(master)        # the server's view (OnTick's resynchronisation); a client's otherwise
        self.n = int(n)
        self.friction = float(friction)
        self.tick_interval = int(tick_interval_ms)
        self.dt = tick_interval_ms * 0.001
        self.fdt = float(torch.tensor(self.dt, dtype=F64).to(torch.float32))
        self.ffriction = float(torch.tensor(self.friction, dtype=F64).to(torch.float32))
        self.dynamical_orientation = bool(dynamical_orientation)
        self.use_new_orbit = bool(use_new_orbit)
        self.block = int(block)
        self.dev = torch.device(device)
        if self.dev.type == "cuda" and self.dev.index is None:
            self.dev = torch.device("cuda", torch.cuda.current_device())
        if self.dev.type == "cuda":
            N.hip()                                  # loud failure now if the kernels are not built
        else:
            N.cpu()
        self.stride = max(64, (self.n + 63) // 64 * 64)
        self._f = torch.zeros(N.NF, self.stride, dtype=F64, device=self.dev)
        self._q = torch.zeros(N.NI, self.stride, dtype=I64, device=self.dev)
        self._b = torch.zeros(8, self.stride, dtype=U8, device=self.dev)
        self._sf = torch.zeros(N.NS, self.stride, dtype=F64, device=self.dev)      # the general launch's scratch
        self._sb = torch.zeros(self.stride, dtype=U8, device=self.dev)
        self.collisions = bool(collisions)      # the gradient method (SpaceBallpark's partition and Gradient)
        self._cf = self._cq = None
        self._hits = torch.zeros(1, self.stride, dtype=I64, device=self.dev)
        if self.collisions:
            self._cf = torch.zeros(N.NCF, self.stride, dtype=F64, device=self.dev)
            self._cq = torch.zeros(N.NCI, self.stride, dtype=I64, device=self.dev)
        self._st = None                # the static collidables' table (load_statics): device tensors and the layout
        self.display_ypr = bool(display_ypr)
        self._y = torch.zeros(N.NY, self.stride, dtype=F64, device=self.dev) if self.display_ypr else None
        self.current_time = 0
        self.refiled = torch.zeros(self.n, dtype=I64, device=self.dev)
        self.last_collision = -1.0
        self.follow_range = 10.0
        self.ids = torch.arange(1, self.n + 1, dtype=I64, device=self.dev)
        self.follow_ptr = -1
        self.bubble = -1
        self.max_ang_vel = 1.0
        self.rot_agility = float(f32(torch.tensor(0.9, dtype=F64)))
        self.rot = torch.tensor([0.0, 0.0, 0.0, 1.0], dtype=F64, device=self.dev).expand(self.n, 4)
        self.old_rot = self.rot.clone()
        self._forms = None             # load_formations' table (formations, slots, lengths) on the device
        self._fslots = torch.zeros(self.n, dtype=I64, device=self.dev)   # each ball's reserved formation slots
        self._fld = None               # the force fields' keys [3, stride] (harmonic, corporation, alliance) once set
        self.tiles = None
        self.rowmap = None             # [n] int32: the tiled launch's rows in a regrouped world order (None: row order)
        self._state = None
        self.schedule = 0              # lanes of a tile: 0 row order, 1 the built-in branch key, 2 `keys`, 3 the
                                       # rows' last search steps (`iters`, record_paths)
        self.keys = None               # [n] uint8 buckets for schedule 2 (the learned scheduler's)
        self.path = None               # [n] uint8: each row's executed path class, when recorded
        self.iters = None              # [n] uint8: each row's orientation searches' steps, when recorded
        self._set_tiles()

    # ---- the state as the kernel takes it ----

    def _set_tiles(self) -> None:
        """Tiles of whole worlds for the tiled launch when every world's rows are contiguous and fit a workgroup;
        None (the general launch) otherwise. Restores the row order (rowmap None)."""
        self.tiles = None
        self.rowmap = None
        self._state = None
        n = self.n
        if n == 0:
            return
        w = self.world.cpu().numpy()
        starts = np.concatenate(([0], np.flatnonzero(w[1:] != w[:-1]) + 1))
        sizes = np.diff(np.concatenate((starts, [n])))
        if len(np.unique(w[starts])) != len(starts) or int(sizes.max()) > self.block:
            return
        tiles = [0]
        for s0, size in zip(starts.tolist(), sizes.tolist()):
            if s0 + size - tiles[-1] > self.block:
                tiles.append(s0)
        tiles.append(n)
        self.tiles = torch.tensor(tiles, dtype=torch.int32, device=self.dev)

    def state(self) -> N.ModesState:
        """The C ABI's xb_modes_state for this park (its blocks never move, so it is made once)."""
        if self._state is None:
            s = N.ModesState()
            s.f64, s.i64, s.u8 = self._f.data_ptr(), self._q.data_ptr(), self._b.data_ptr()
            s.scratch_f64, s.scratch_u8 = self._sf.data_ptr(), self._sb.data_ptr()
            s.n, s.stride = self.n, self.stride
            s.new_orbit = int(self.use_new_orbit)
            s.friction, s.dt, s.ffriction, s.fdt = self.friction, self.dt, self.ffriction, self.fdt
            s.ticks = 1
            self._state = s
        s = self._state
        s.new_orbit = int(self.use_new_orbit)
        s.dynamical = int(self.dynamical_orientation)
        tiled = self.tiles is not None
        s.tiles = self.tiles.data_ptr() if tiled else None
        s.n_tiles = (len(self.tiles) - 1) if tiled else 0
        s.rowmap = self.rowmap.data_ptr() if (tiled and self.rowmap is not None) else None
        s.schedule = int(self.schedule) if tiled else 0
        s.n_keys = 16
        s.keys = self.keys.data_ptr() if (self.keys is not None and s.schedule == 2) else None
        s.path = self.path.data_ptr() if self.path is not None else None
        s.iters = self.iters.data_ptr() if self.iters is not None else None
        s.collisions = int(self.collisions)
        if self.collisions:
            s.coll_f64, s.coll_i64 = self._cf.data_ptr(), self._cq.data_ptr()
            s.hits, s.hits_cap = self._hits.data_ptr(), self._hits.shape[0]
        s.ypr = self._y.data_ptr() if self.display_ypr else None
        st = self._st if self.collisions else None
        s.statics = int(st is not None)
        s.st_f64 = st["f"].data_ptr() if st else None
        s.st_i64 = st["q"].data_ptr() if st else None
        s.st_keys = st["keys"].data_ptr() if st else None
        s.st_range = st["range"].data_ptr() if st else None
        s.st_stride = st["stride"] if st else 0
        s.st_boxes = st["boxes"] if st else 0
        fm = self._forms
        s.form = fm["table"].data_ptr() if fm else None
        s.form_len = fm["len"].data_ptr() if fm else None
        s.n_form = fm["n"] if fm else 0
        s.form_stride = fm["stride"] if fm else 0
        s.fld = self._fld.data_ptr() if self._fld is not None else None
        s.form_slots = self._fslots.data_ptr() if (fm and self.master) else None
        return s

    def record_paths(self, on: bool = True) -> None:
        """Record at every tick each row's executed path class (XB_PATH_*: 0 idle .. 7 warp edge) in `path` and its
        orientation searches' steps (GotoTorqueBreaks' and GotoRollBreaks' inner iterations, at most 255) in `iters`:
        what its tick cost, never a result."""
        self.path = torch.zeros(self.n, dtype=U8, device=self.dev) if on else None
        self.iters = torch.zeros(self.n, dtype=U8, device=self.dev) if on else None
        self._state = None

    def row_tensors(self) -> list:
        """The per-row state (the reference's `row_tensors`, vec_orient.py:814-818): every field of the blocks."""
        return list(FIELD_NAMES)

    # ---- loading: the Python AddBall of new balls (vec_modes.py:221-282) ----

    def _decay(self, m: torch.Tensor) -> torch.Tensor:
        """SetBallTimeFactor: 0 if mass x agility <= 0, else glibc's exp(-k dt / m) (NaN for a NaN m)."""
        x = torch.full_like(m, -self.friction * self.dt) / m
        return torch.where(m <= 0.0, 0.0, N.maths("exp", x))

    def load(self, rows: list) -> None:
        """Every ball as the Python AddBall adds a new one (vec_modes.py:221-282): keys id, world, time (the world's
        clock), mass, radius, max_vel, pos, vel, agility, speed_fraction, free, massive, interactive, cloaked, bubble,
        moribund, mode; with dynamical orientation also max_ang_vel (default 1) and rot_agility (default 0.9)."""
        n = len(rows)
        if n != self.n:
            raise ValueError(f"{n} rows for a park of {self.n}")
        t = lambda xs, dt=F64: torch.tensor(xs, dtype=dt)
        g = lambda key, d: [r.get(key, d) for r in rows]
        self.load_arrays(
            ids=t([int(r["id"]) for r in rows], I64), world=t([int(x) for x in g("world", 0)], I64),
            time=t([int(x) for x in g("time", 0)], I64), mass=t([float(r["mass"]) for r in rows]),
            radius=t([float(r["radius"]) for r in rows]), max_vel=t([float(r["max_vel"]) for r in rows]),
            pos=t([[float(c) for c in r["pos"]] for r in rows]).reshape(n, 3),
            vel=t([[float(c) for c in r["vel"]] for r in rows]).reshape(n, 3),
            agility=t([float(r["agility"]) for r in rows]),
            speed_fraction=t([float(r["speed_fraction"]) for r in rows]),
            free=t([bool(x) for x in g("free", True)], torch.bool),
            massive=t([bool(x) for x in g("massive", False)], torch.bool),
            interactive=t([bool(x) for x in g("interactive", False)], torch.bool),
            cloaked=t([int(x) for x in g("cloaked", 0)], I64), bubble=t([int(x) for x in g("bubble", -1)], I64),
            moribund=t([bool(x) for x in g("moribund", False)], torch.bool),
            mode=t([int(x) for x in g("mode", STOP)], I64),
            max_ang_vel=t([float(x) for x in g("max_ang_vel", 1.0)]),
            rot_agility=t([float(x) for x in g("rot_agility", 0.9)]),
            junk=t([bool(x) for x in g("junk", False)], torch.bool))
        if any("owner" in r for r in rows):             # a miniball's parent (mini_rows)
            self.owner_id = t([int(x) for x in g("owner", 0)], I64).to(self.dev)

    def load_arrays(self, *, ids, world, time, mass, radius, max_vel, pos, vel, agility, speed_fraction, free,
                    massive, interactive, cloaked, bubble, moribund, mode, max_ang_vel=None, rot_agility=None,
                    junk=None) -> None:
        """`load` from [n] and [n, 3] tensors: AddBall's clamps (the floats rounded to float, then `x < 0 ? 0 : x`,
        agility `<= 0 ? 1`, NaN and -0 kept), the decay factor, Ball::Ball's defaults (follow range 10, warp length
        -1, no follow fields, the sensor off, the turning values 1 rad/s and 0.9). With dynamical orientation
        AddDynamicallyOrientedBall's turning values (max_ang_vel `< 0 ? 0`, rot_agility `<= 0 ? 1`, as floats; default
        1 and 0.9), the identity rotation at rest, its old rotation, and the rotation snapped to the velocity
        (src/exact_ballpark/vec_modes.py:252-264)."""
        dev = self.dev
        c = lambda x, dt=F64: torch.as_tensor(x).to(device=dev, dtype=dt)
        if int(((c(mode, I64) != STOP) & ~torch.isin(c(mode, I64), torch.tensor(INERT_MODES, device=dev))).sum()):
            raise ValueError("a ball starts in STOP or an inert mode")
        m = c(mass)
        self.mass = torch.where(m < 0.0, 0.0, m)
        x = f32(c(radius))
        self.radius = torch.where(x < 0.0, 0.0, x)
        x = f32(c(max_vel))
        self.max_vel = torch.where(x < 0.0, 0.0, x)
        x = f32(c(agility))
        mini = c(mode, I64) == MINIBALL                 # AddMiniball's own values, without AddBall's clamps
        self.agility = torch.where((x <= 0.0) & ~mini, 1.0, x)
        x = f32(c(speed_fraction))
        self.speed_fraction = torch.where(x < 0.0, 0.0, x)
        p, v = c(pos), c(vel)
        self.pos, self.vel = p, v
        self.old_pos = p - v * self.dt
        self.old_vel = v
        self.goto, self.last_g = 0.0, 0.0
        self.time_factor = self._decay(self.mass * self.agility)
        self.rot = torch.tensor([0.0, 0.0, 0.0, 1.0], dtype=F64, device=dev).expand(self.n, 4)
        self.old_rot = self.rot.clone()
        self.ang_vel, self.old_ang_vel, self.torque, self.rot_roll = 0.0, 0.0, 0.0, 0.0
        if self.dynamical_orientation:
            x = f32(c(1.0 if max_ang_vel is None else max_ang_vel).expand(self.n))
            self.max_ang_vel = torch.where(x < 0.0, 0.0, x)
            x = f32(c(0.9 if rot_agility is None else rot_agility).expand(self.n))
            self.rot_agility = torch.where(x <= 0.0, 1.0, x)
        else:
            self.max_ang_vel = 1.0
            self.rot_agility = float(f32(torch.tensor(0.9, dtype=F64)))
        self.ids, self.world, self.clock, self.mode = c(ids, I64), c(world, I64), c(time, I64), c(mode, I64)
        self._index = None                            # lookup's (world, id) -> row map, rebuilt on demand
        self.free, self.massive = c(free, torch.bool), c(massive, torch.bool)
        self.interactive, self.moribund = c(interactive, torch.bool), c(moribund, torch.bool)
        self.cloaked, self.bubble = c(cloaked, I64), c(bubble, I64)
        self.sensor = False
        self.effect_stamp, self.follow_id, self.owner_id = 0, 0, 0
        self.follow_ptr = -1
        self.follow_range = 10.0
        self.last_collision = -1.0
        self.junk = False if junk is None else c(junk, torch.bool)
        self.formation_id = -1
        self._dying = torch.zeros(self.n, dtype=torch.bool, device=dev)    # removed with a delay: the moribund set
        self._mush = torch.zeros(self.n, dtype=torch.bool, device=dev)     # mushrooms (in the set once removed)
        empty = self.ids == EMPTY_ID
        if bool((empty & (self.free | self.massive | ~self.moribund)).any()):
            raise ValueError("an empty row (id EMPTY_ID) is moribund, fixed and not massive: use empty_row()")
        self._set_tiles()
        self._fld = None
        self._fslots.zero_()
        if self.dynamical_orientation:                  # a miniball keeps the identity, old and new (full.py:1350)
            self._orient_setup(None if not bool(mini.any()) else ~mini, 1)
        if self.display_ypr:                            # AddBall's snap of a free ball (full.py:1283-1288)
            self._y.zero_()
            self._ypr_snap(self.free)
        if self.collisions:
            self._collide_setup()

    # ---- the gradient collision method's state (space.py's SpaceBallpark, a client's view) ----

    def _collide_setup(self) -> None:
        """After a load: the collision lists' room (3 (k - 1) + 1 entries for a largest world of k rows: Potential
        appends at most three entries a candidate, one per recursion depth), mLastC and the lists cleared, every ball
        filed (AddBall's InsertBallInBoxes)."""
        if self.n == 0:
            return
        w = self.world.cpu().numpy()
        starts = np.concatenate(([0], np.flatnonzero(w[1:] != w[:-1]) + 1))
        biggest = int(np.diff(np.concatenate((starts, [self.n]))).max())
        if self.tiles is None or biggest > MAX_COLLIDE_WORLD:
            raise ValueError(f"collisions need each world's rows consecutive and at most {MAX_COLLIDE_WORLD} rows a "
                             f"world, within a workgroup of {self.block}")
        cap = 3 * (biggest - 1) + 1 + (self._st["most"] if self._st else 0)     # and one entry a static collidable
        self._hits = torch.full((cap, self.stride), -1, dtype=I64, device=self.dev)
        if self._st:
            self._st_set_range()
        self._cf.zero_()
        self._cq.zero_()
        self._state = None
        self._refile(None)

    # ---- static collidables (FullBallpark's capsules and oriented boxes; VecCollidePark's load_statics) ----

    def load_statics(self, statics) -> None:
        """Each world's static collidables, as AddCapsule and AddOrientedBox add them (a client's view), filed once:
        `statics` a list indexed by world id, or a dict {world id: list}, of {"kind": "capsule", "id", "parent", "a",
        "b", "radius"} (the radius a float) or {"kind": "box", "id", "parent", "corner", "lx", "ly", "lz"}; ids
        distinct within a world, at most 64 a world (src/exact_ballpark/vec_collide.py:1329-1360). A park with
        collisions only; they take part in every tick's gradient from now on (an empty `statics`: none)."""
        items = statics.items() if isinstance(statics, dict) else enumerate(statics)
        world, kind, ids, parents, geom = [], [], [], [], []
        for w, recs in items:
            for s in recs:
                if s["kind"] == "capsule":
                    g = [float(x) for x in s["a"]] + [float(x) for x in s["b"]] + [float(s["radius"])] + [0.0] * 5
                    kind.append(0)
                elif s["kind"] == "box":
                    g = [float(x) for k in ("corner", "lx", "ly", "lz") for x in s[k]]
                    kind.append(1)
                else:
                    raise ValueError(f"unknown static collidable kind {s['kind']!r}")
                if len(g) != 12:
                    raise ValueError(f"static collidable {s['id']}: three coordinates a point or vector")
                world.append(int(w))
                ids.append(int(s["id"]))
                parents.append(int(s["parent"]))
                geom.append(g)
        self.load_statics_arrays(world=world, kind=kind, id=ids, parent=parents, geom=np.reshape(geom, (-1, 12)))

    def load_statics_arrays(self, *, world, kind, id, parent, geom) -> None:
        """`load_statics` from arrays: [S] world ids, kinds (0 capsule, 1 box), ids and parent ids, and [S, 12]
        arguments (a capsule's A, B and radius, then zeros; a box's corner, lx, ly, lz); any order of worlds."""
        if not self.collisions:
            raise ValueError("static collidables take part in the gradient method: make the park with collisions=True")
        world, kind = np.asarray(world, dtype=np.int64), np.asarray(kind, dtype=np.int64)
        ids, par = np.asarray(id, dtype=np.int64), np.asarray(parent, dtype=np.int64)
        geom = np.ascontiguousarray(np.asarray(geom, dtype=np.float64).reshape(-1, 12))
        n = len(world)
        if not (len(kind) == len(ids) == len(par) == len(geom) == n):
            raise ValueError("static collidables: the arrays' lengths differ")
        if n == 0:
            self._st = None
            self._collide_setup()
            return
        if ((kind != 0) & (kind != 1)).any():
            raise ValueError(f"unknown static collidable kind {int(kind[(kind != 0) & (kind != 1)][0])}")
        order = np.argsort(world, kind="stable")
        world, kind, ids, par, geom = world[order], kind[order], ids[order], par[order], geom[order]
        wid, first, count = np.unique(world, return_index=True, return_counts=True)
        if int(count.max()) > MAX_WORLD_STATICS:
            raise ValueError(f"a world has {int(count.max())} static collidables, at most {MAX_WORLD_STATICS}")
        key = np.stack((world, ids), axis=1)
        if len(np.unique(key, axis=0)) != n:
            raise ValueError("static collidable ids must be distinct within a world")
        stride = max(64, (n + 63) // 64 * 64)
        f = np.zeros((N.NSF, stride), dtype=np.float64)
        q = np.zeros((N.NSI, stride), dtype=np.int64)
        boxes = 16
        while True:
            keys = np.zeros((3 * boxes, stride), dtype=np.int64)
            most = N.cpu().xb_cpu_statics_setup(n, kind.ctypes.data, ids.ctypes.data, par.ctypes.data,
                                                  geom.ctypes.data, f.ctypes.data, q.ctypes.data, keys.ctypes.data,
                                                  stride, boxes)
            if most == -2:
                raise ValueError("Capsule::InsertInBoxes never ends (twice the radius is below half a float ulp of the "
                                 "length left; Capsule.cpp:499, 578): Destiny's result is undefined")
            if most < 0:
                raise ValueError("xb_cpu_statics_setup refused the static collidables")
            if most <= boxes:
                break
            boxes = most
        t = lambda a: torch.from_numpy(a).to(self.dev)
        self._st = {"f": t(f), "q": t(q), "keys": t(keys), "stride": stride, "boxes": boxes, "wid": wid,
                    "first": first.astype(np.int32), "count": count.astype(np.int32), "most": int(count.max()),
                    "range": None}
        self._collide_setup()

    def _st_set_range(self) -> None:
        """Each row's world's first static collidable in the table and how many."""
        w = self.world.cpu().numpy()
        st = self._st
        at = np.clip(np.searchsorted(st["wid"], w), 0, len(st["wid"]) - 1)
        has = st["wid"][at] == w
        rng = np.zeros((2, self.stride), dtype=np.int32)
        rng[0, :self.n] = np.where(has, st["first"][at], 0)
        rng[1, :self.n] = np.where(has, st["count"][at], 0)
        st["range"] = torch.from_numpy(rng).to(self.dev)
        self._state = None

    def static_boxes(self, world: int, k: int) -> np.ndarray:
        """[boxes, 3]: the keys world `world`'s k-th static collidable is filed in (at its level), sorted."""
        i = int(self._st["first"][int(np.searchsorted(self._st["wid"], int(world)))]) + int(k)
        nb = int(self._st["q"][4, i])
        keys = self._st["keys"][:3 * nb, i].cpu().numpy().reshape(nb, 3)
        return keys[np.lexsort(keys.T[::-1])]

    def _refile(self, mask: torch.Tensor | None) -> None:
        """InsertBallInBoxes for the rows in mask (None: every row): their filing from their current state."""
        if not self.collisions or self.n == 0:
            return
        rows = None
        if mask is not None:
            rows = torch.nonzero(mask.to(self.dev).reshape(self.n)).reshape(-1).to(torch.int32)
            if rows.numel() == 0:
                return
        nr = self.n if rows is None else rows.numel()
        s = self.state()
        ptr = None if rows is None else rows.data_ptr()
        if self.dev.type == "cuda":
            with torch.cuda.device(self.dev):
                rc = N.hip().xb_collide_file(ctypes.byref(s), ptr, nr, torch.cuda.current_stream(self.dev).cuda_stream)
            N.check(rc, "xb_collide_file")
        else:
            N.check(N.cpu().xb_cpu_collide_file(ctypes.byref(s), ptr, nr), "xb_cpu_collide_file")

    def _ypr_snap(self, mask: torch.Tensor | None) -> None:
        """xb_modes_ypr_snap on the rows in mask (None: every row): CalculateYawPitchRoll with snap."""
        if not self.display_ypr or self.n == 0:
            return
        rows = None
        if mask is not None:
            rows = torch.nonzero(mask.to(self.dev).reshape(self.n)).reshape(-1).to(torch.int32)
            if rows.numel() == 0:
                return
        nr = self.n if rows is None else rows.numel()
        s = self.state()
        ptr = None if rows is None else rows.data_ptr()
        if self.dev.type == "cuda":
            with torch.cuda.device(self.dev):
                rc = N.hip().xb_modes_ypr_snap(ctypes.byref(s), ptr, nr,
                                               torch.cuda.current_stream(self.dev).cuda_stream)
            N.check(rc, "xb_modes_ypr_snap")
        else:
            N.check(N.cpu().xb_cpu_modes_ypr_snap(ctypes.byref(s), ptr, nr), "xb_cpu_modes_ypr_snap")

    @property
    def hits(self) -> torch.Tensor:
        """[n, cap] int64: each row's collision list of its last tick (the other balls' ids in Potential's order), its
        first hits_n entries; entries past them are left from earlier ticks."""
        if not self.collisions:
            raise AttributeError("hits: the park was made without collisions")
        return self._hits[:, :self.n].t()

    def collision_lists(self) -> list:
        """Each row's collision list of its last tick, as Python lists (the reference's `collisions`)."""
        n = self.hits_n.cpu().tolist()
        h = self.hits.cpu().numpy()
        if max(n, default=0) > h.shape[1]:
            raise RuntimeError("a collision list past its room")
        return [[int(x) for x in h[r, :k]] for r, k in enumerate(n)]

    # ---- commands and setters: one record a listed row, run by xb_modes_command_apply ----

    def _command(self, op: str, mask: torch.Tensor, *, flag=None, ints=(), x=None, check: bool = False,
                 release: bool = False) -> torch.Tensor | None:
        """Runs command `op` (CMD) on the rows in `mask` [n] in one batch: each row's flag [n], integer arguments
        (tensors [n]) and float arguments x [n, k] into its record (include/xb.h's xb_command); `check`: the gather
        step first, its refusals raised before anything changes; `release`: StopAllFollowers for the rows the
        command marks. Returns the records (None when no row is listed)."""
        if self.n == 0:
            return None
        d = self.dev
        rows = torch.nonzero(mask.to(d).reshape(self.n)).reshape(-1)
        m = rows.numel()
        if m == 0:
            return None
        rec = torch.zeros(m, CMD_WORDS, dtype=I64, device=d)
        rec[:, 0] = CMD[op]
        if flag is not None:
            rec[:, 1] = flag.to(device=d, dtype=I64).reshape(self.n)[rows]
        for k, v in enumerate(ints):
            rec[:, 2 + k] = v.to(device=d, dtype=I64).reshape(self.n)[rows]
        if x is not None:
            xv = x.to(device=d, dtype=F64).reshape(self.n, -1)[rows]
            rec.view(F64)[:, 5:5 + xv.shape[1]] = xv
        rows = rows.to(torch.int32)
        s = self.state()
        mark = None
        if release:
            if getattr(self, "_mark", None) is None:
                self._mark = torch.zeros(self.stride, dtype=U8, device=d)
            mark = self._mark.data_ptr()
        if d.type == "cuda":
            with torch.cuda.device(d):
                stream = torch.cuda.current_stream(d).cuda_stream
                if check:
                    status = torch.zeros(1, dtype=torch.int32, device=d)
                    N.check(N.hip().xb_modes_command_check(ctypes.byref(s), rows.data_ptr(), m, rec.data_ptr(),
                                                           status.data_ptr(), stream), "xb_modes_command_check")
                    _refusal(int(status.item()))
                N.check(N.hip().xb_modes_command_apply(ctypes.byref(s), rows.data_ptr(), m, rec.data_ptr(), mark,
                                                       stream), "xb_modes_command_apply")
        else:
            if check:
                status = ctypes.c_int32(0)
                N.check(N.cpu().xb_cpu_modes_command_check(ctypes.byref(s), rows.data_ptr(), m, rec.data_ptr(),
                                                           ctypes.byref(status)), "xb_cpu_modes_command_check")
                _refusal(status.value)
            N.check(N.cpu().xb_cpu_modes_command_apply(ctypes.byref(s), rows.data_ptr(), m, rec.data_ptr(), mark),
                    "xb_cpu_modes_command_apply")
        return rec

    def _stop_rows(self, mask: torch.Tensor) -> None:
        """Stop(Ball*) on the rows in mask: a follower forgets its target; the mode becomes STOP."""
        self._command("stop", mask)

    def _release_followers(self, mask: torch.Tensor) -> None:
        """StopAllFollowers of the balls in mask: their missiles (no longer massive) and interactive orbiters fly on
        along their velocities, their other followers stop."""
        self._command("release", mask, release=True)

    def followers(self, row: int) -> list:
        """The ids of the balls following the ball of `row` (their follow pointer at the row, naming it by id)."""
        fol = torch.nonzero((self.follow_ptr == int(row)) & (self.follow_id == self.ids[int(row)])).reshape(-1)
        return sorted(int(i) for i in self.ids[fol].tolist())

    def stop(self, mask: torch.Tensor) -> None:
        self._command("stop", mask)

    def goto_point(self, mask: torch.Tensor, point: torch.Tensor) -> None:
        self._command("goto", mask, x=point)

    def goto_direction(self, mask: torch.Tensor, direction: torch.Tensor) -> None:
        self._command("goto_direction", mask, x=direction)

    def _refuse_mushroom_targets(self, ok: torch.Tensor, trow: torch.Tensor) -> None:
        """A mushroom as a target (its followers would be released at its removal inside the tick, before or after
        their own turns): not batched, an error."""
        if bool((ok & (self.mode[trow] == MUSHROOM)).any()):
            raise ValueError("a mushroom as a follow, orbit, missile or formation target is not supported")

    def follow(self, mask: torch.Tensor, target: torch.Tensor, range_: torch.Tensor) -> None:
        self._command("follow", mask, ints=(target,), x=range_, check=True)

    def orbit(self, mask: torch.Tensor, target: torch.Tensor, range_: torch.Tensor) -> None:
        self._command("orbit", mask, ints=(target,), x=range_, check=True)

    # ---- missiles (VecCombatPark's API, src/exact_ballpark/vec_combat.py:339-396; full.py:1679-1736) ----

    def lookup(self, ball: torch.Tensor) -> torch.Tensor:
        """[n] the row of Destiny id ball[row] in row's world, or -1 (VecCombatPark.lookup, vec_combat.py:217-223)."""
        if getattr(self, "_index", None) is None:
            w, ids = self.world.cpu().tolist(), self.ids.cpu().tolist()
            self._index = {(a, b): r for r, (a, b) in enumerate(zip(w, ids))}
        w = self.world.cpu().tolist()
        want = ball.to(I64).cpu().tolist()
        rows = [self._index.get((a, b), -1) for a, b in zip(w, want)]
        return torch.tensor(rows, dtype=I64, device=self.dev)

    def missile_follow(self, mask: torch.Tensor, target: torch.Tensor, owner: torch.Tensor) -> None:
        """MissileFollow for the missiles in `mask` (target and owner Destiny ids; a negative owner: a defender): refused
        for a negative target, the missile, its owner or a moribund target as target; stopped; then, unless the target is
        in another bubble or cloaked, aimed at it (follow range minus both radii, a float), straight on for 800 ms
        toward 1e16 m along its velocity, at full speed fraction."""
        dst = target.to(device=self.dev, dtype=I64)
        self._command("missile_follow", mask, ints=(self.lookup(dst), dst, owner), check=True)

    def launch_missile(self, mask: torch.Tensor, target: torch.Tensor, owner: torch.Tensor, aimed: torch.Tensor,
                       massive: torch.Tensor) -> None:
        """LaunchMissile for the missile balls in `mask` (target and owner Destiny ids, the launcher |owner|; aimed and
        massive the thunk's chars): from the launcher's position, aimed at the target (unit speed) or along the
        launcher's velocity (its heading backwards when at rest) at max(150, its maximum speed); massive if asked;
        SetBallPosition (re-filed) and SetBallVelocity (snapped); then MissileFollow, or GotoDirection without a
        target. At rest without dynamical orientation the heading is the display orientation's (display_ypr)."""
        dst = target.to(device=self.dev, dtype=I64)
        own = owner.to(device=self.dev, dtype=I64)
        launcher = torch.where(dst >= 0, self.lookup(own.abs()), -1)
        tgt = torch.where(dst != own, self.lookup(dst), -1)
        flag = (aimed.to(device=self.dev, dtype=I64) & 0xFF) | ((massive.to(device=self.dev, dtype=I64) & 0xFF) << 8)
        self._command("launch_missile", mask.to(self.dev) & (launcher >= 0), flag=flag, ints=(launcher, tgt, own),
                      check=True)

    # ---- removal ----

    def remove(self, mask: torch.Tensor, delay: torch.Tensor | None = None) -> None:
        """RemoveBall between ticks (full.py:1633-1675; VecCombatPark.remove, vec_combat.py:520), a client's view,
        for the balls in `mask`: stopped, their followers released (missiles and interactive orbiters fly on); without
        a delay their rows emptied (no ball: the tick skips them, no ball meets them; add_mushroom may take them);
        with a delay of d > 0 ticks moribund and non-massive, stamped with the clock plus d, until
        bring_out_the_deads removes them. A fixed ball's minis are not removed with it: refused with an error."""
        mask = mask.to(self.dev) & (self.ids != EMPTY_ID)
        if not bool(mask.any()):
            return
        if bool((mask & ~self.free & (torch.isin(self.ids, self.owner_id[self.mode == MINIBALL]))).any()):
            raise ValueError("remove: a parent with minis (its minis' removal is not batched)")
        if delay is not None:
            d = delay.to(device=self.dev, dtype=I64)
            later = mask & (d > 0)
            if bool(later.any()):
                self._command("stop_release", later, release=True)
                self.moribund = torch.where(later, True, self.moribund)
                self.massive = torch.where(later, False, self.massive)
                self.effect_stamp = torch.where(later, self.clock + d, self.effect_stamp)
                self._dying = self._dying | later
                mask = mask & ~later
                if not bool(mask.any()):
                    return
        self._command("stop_release", mask, release=True)
        self.ids = torch.where(mask, EMPTY_ID, self.ids)
        self._index = None
        self.moribund = torch.where(mask, True, self.moribund)
        self.massive = torch.where(mask, False, self.massive)
        self.free = torch.where(mask, False, self.free)
        self.mode = torch.where(mask, STOP, self.mode)
        self._fslots.masked_fill_(mask, 0)              # the ball is gone with its reservations
        self._dying = self._dying & ~mask
        self._state = None

    def bring_out_the_deads(self, worlds=None) -> None:
        """BringOutTheDeads between ticks (space.py:1319-1341), as OnTick calls it, in the given worlds (default
        every world): over each world's moribund set (the balls removed with a delay, a mushroom removed in a tick)
        by ascending id, a ball no longer moribund leaves the set, one whose stamp has passed is removed, and up to 7
        more within 2 ticks of theirs."""
        dying = self._dying | (self._mush & self.moribund & (self.ids != EMPTY_ID))
        if not bool(dying.any()):
            return
        rows = torch.nonzero(dying).reshape(-1).cpu().tolist()
        world, ids, stamp, clock, mor = (x.cpu().tolist() for x in (self.world, self.ids, self.effect_stamp,
                                                                    self.clock, self.moribund))
        keep = None if worlds is None else set(int(w) for w in worlds)
        by_world: dict = {}
        for r in rows:
            if keep is None or world[r] in keep:
                by_world.setdefault(world[r], []).append(r)
        due, left = [], []
        for w, rs in by_world.items():
            kill = 0
            for r in sorted(rs, key=lambda r: ids[r]):
                if not mor[r]:
                    left.append(r)
                elif stamp[r] - clock[r] < 0:
                    due.append(r)
                elif stamp[r] - clock[r] - 2 < 0 and kill < 7:
                    due.append(r)
                    kill += 1
        if left:
            out = torch.zeros(self.n, dtype=torch.bool, device=self.dev)
            out[left] = True
            self._dying = self._dying & ~out
            self._mush = self._mush & ~out
        if due:
            m = torch.zeros(self.n, dtype=torch.bool, device=self.dev)
            m[due] = True
            self._mush = self._mush & ~m
            self.remove(m)

    # ---- collision events ----

    def collision_events(self) -> list:
        """Ball::DispatchCollisions of the last tick (full.py:2571-2583; COL-719), a client's view: for each live
        missile that moved at least sqrt(0.1) m in it, a DoCollision event per entry of its collision list, as
        (world, time, missile id, collider id, mLastC x, y, z), worlds and balls by ascending id, each list in its
        order. Read after each one-tick launch: a launch of several ticks leaves its last tick's only."""
        if not self.collisions or self.n == 0:
            return []
        d = self.pos - self.old_pos
        sq = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        sel = self.free & ~self.moribund & (self.mode == MISSILE) & ~(sq < 0.1) & (self.hits_n > 0)
        ticked = getattr(self, "_ticked", None)
        if ticked is not None and ticked[0] == self.current_time:
            sel = sel & ticked[1]                  # a world the last (masked) tick left out posts no event
        rows = torch.nonzero(sel).reshape(-1)
        if rows.numel() == 0:
            return []
        r = rows.cpu().tolist()
        world, ids, clock = (x[rows].cpu().tolist() for x in (self.world, self.ids, self.clock))
        nh = self.hits_n[rows].cpu().tolist()
        lc = self.last_c[rows].cpu().tolist()
        hits = self._hits[:, rows].cpu().tolist()
        out = []
        for k in sorted(range(len(r)), key=lambda k: (world[k], ids[k])):
            for h in range(nh[k]):
                out.append((world[k], clock[k] - 1, ids[k], hits[h][k], *lc[k]))
        return out

    # ---- minis and force fields (VecCollidePark's, src/exact_ballpark/vec_collide.py:364-420, 748-770) ----

    @staticmethod
    def mini_rows(balls: list, minis: list, world: int, time: int = 0) -> tuple:
        """A world's minis as a client's Destiny makes them for fixed parents at the load (make_minis,
        vec_collide.py:372; full.py:1336-1357, 1476-1536), each taking the next local id (LOCAL_BALLS - 1, - 2, ... in
        the order of `minis`; a free or missing parent's take none): (the miniballs' `load` rows, the minicapsules'
        and miniboxes' `load_statics` records). `balls` are the world's `load` rows, `minis` records {"kind": "ball",
        "parent", "offset", "radius"}, {"kind": "capsule", "parent", "a", "b", "radius"} (offsets from the parent)
        or {"kind": "box", "parent", "corner", "lx", "ly", "lz"} (the corner an offset). A miniball: fixed, massive,
        not interactive, MINIBALL, the parent's mass, the radius a float, maximum speed, agility and speed fraction 0,
        at the parent's position plus its offset, owned by the parent."""
        by_id = {int(b["id"]): b for b in balls}
        local, rows, statics = LOCAL_BALLS, [], []
        for m in minis:
            par = by_id.get(int(m["parent"]))
            if par is None or par.get("free", True):
                continue
            local -= 1
            pp = [float(c) for c in par["pos"]]
            add = lambda o: [pp[i] + float(o[i]) for i in range(3)]
            if m["kind"] == "ball":
                rows.append({"id": local, "world": world, "time": time, "mass": float(par["mass"]),
                             "radius": float(m["radius"]), "max_vel": 0.0, "pos": add(m["offset"]),
                             "vel": [0.0, 0.0, 0.0], "agility": 0.0, "speed_fraction": 0.0, "free": False,
                             "massive": True, "interactive": False, "cloaked": 0, "bubble": -1, "moribund": False,
                             "mode": MINIBALL, "owner": int(par["id"])})
            elif m["kind"] == "capsule":
                statics.append({"kind": "capsule", "id": local, "parent": int(par["id"]), "a": add(m["a"]),
                                "b": add(m["b"]), "radius": float(m["radius"])})
            else:
                statics.append({"kind": "box", "id": local, "parent": int(par["id"]), "corner": add(m["corner"]),
                                "lx": list(m["lx"]), "ly": list(m["ly"]), "lz": list(m["lz"])})
        return rows, statics

    def set_harmonic(self, mask: torch.Tensor, harmonic: torch.Tensor, corporation: torch.Tensor,
                     alliance: torch.Tensor, is_field: torch.Tensor) -> None:
        """SetBallHarmonic (space.py:1272-1282) for the balls in `mask`: the harmonic, corporation and alliance; a
        force field is stopped and set to FIELD; clearing a field stops it. The first call gives the park its keys
        (-1 for every ball before)."""
        mask = mask.to(self.dev)
        if self._fld is None:
            self._fld = torch.full((3, self.stride), -1, dtype=I64, device=self.dev)
            self._state = None
        for k, v in enumerate((harmonic, corporation, alliance)):
            self._fld[k, :self.n] = torch.where(mask, v.to(device=self.dev, dtype=I64), self._fld[k, :self.n])
        field = is_field.to(self.dev) != 0
        self._stop_rows(mask & (field | (self.mode == FIELD)))
        self.mode = torch.where(mask & field, FIELD, self.mode)

    @property
    def harmonic(self) -> torch.Tensor:
        return self._fld[0, :self.n] if self._fld is not None else torch.full((self.n,), -1, dtype=I64,
                                                                               device=self.dev)

    # ---- mushrooms (VecCombatPark's API, src/exact_ballpark/vec_combat.py:540-613; full.py:1826-1846) ----

    @staticmethod
    def empty_row(world: int, time: int = 0) -> dict:
        """A `load` row kept empty for a ball added mid-run (add_mushroom): no ball (id EMPTY_ID), moribund, fixed,
        not massive, at the origin; the tick skips it and no ball meets it."""
        return {"id": EMPTY_ID, "world": world, "time": time, "mass": 0.0, "radius": 0.0, "max_vel": 0.0,
                "pos": [0.0, 0.0, 0.0], "vel": [0.0, 0.0, 0.0], "agility": 1.0, "speed_fraction": 1.0, "free": False,
                "massive": False, "interactive": False, "cloaked": 0, "bubble": -1, "moribund": True, "mode": STOP}

    def add_balls(self, mask: torch.Tensor, rows: list) -> None:
        """The Python AddBall of new balls mid-run into the empty rows in `mask` (VecCombatPark.add_balls,
        vec_combat.py:225-266; PyAddBall, Thunkers.cpp:150): one dict a masked row, in row order, with `load`'s keys
        but the world and the clock, which are the row's. Each row is made as `load` makes a ball (AddBall's clamps and
        defaults, the decay factor, STOP, with dynamical orientation the rotation snapped, the display orientation of a
        free ball snapped) in a park of the new balls alone, and its blocks' columns copied in; a collision park files
        it. A row that is not empty, or an id its world already holds, is refused before anything changes."""
        idx = torch.nonzero(mask.to(self.dev).reshape(self.n)).reshape(-1)
        k = idx.numel()
        if k != len(rows):
            raise ValueError(f"add_balls: {len(rows)} balls for {k} rows")
        if k == 0:
            return
        if bool((self.ids[idx] != EMPTY_ID).any()):
            raise ValueError("add_balls: a row is not empty")
        worlds, clocks = self.world[idx].cpu().tolist(), self.clock[idx].cpu().tolist()
        taken = set(zip(self.world.cpu().tolist(), self.ids.cpu().tolist()))
        for w, r in zip(worlds, rows):
            if (w, int(r["id"])) in taken or int(r["id"]) == EMPTY_ID:
                raise ValueError(f"add_balls: ball {r['id']} is in world {w} already")
            taken.add((w, int(r["id"])))
        sub = VecModesPark(k, self.friction, self.tick_interval, self.dev, self.dynamical_orientation,
                           self.use_new_orbit, block=self.block, display_ypr=self.display_ypr)
        sub.load([{**r, "world": w, "time": t} for r, w, t in zip(rows, worlds, clocks)])
        keep_bubble = not all("bubble" in r for r in rows)     # the empty row's bubble otherwise (its world's)
        bubble = self._q[9, idx].clone()
        self._f[:, idx] = sub._f[:, :k]
        self._q[:, idx] = sub._q[:, :k]
        self._b[:, idx] = sub._b[:, :k]
        if keep_bubble:
            self._q[9, idx] = bubble
        if self.display_ypr:
            self._y[:, idx] = sub._y[:, :k]
        if self._fld is not None:
            self._fld[:, idx] = -1
        self._dying[idx] = False
        self._mush[idx] = False
        self._index = None
        if self.collisions:
            self._cf[:, idx] = 0.0
            self._cq[7, idx] = 0                       # the collision list's length
            self._refile(mask)

    def add_mushroom(self, mask: torch.Tensor, owner: torch.Tensor, range_: torch.Tensor, life: torch.Tensor) -> None:
        """AddMushroom into the empty rows in `mask`, a client's view (owner a Destiny id in the row's world, the range
        a float, the lifetime a double): refused (the row left empty) for a range or lifetime not above 0 or a missing
        owner; else ball 0 as AddBall adds it (mass 1e20f, radius 0, maximum speed 0, fixed, massive, not
        interactive, at the owner's position, at rest, agility and speed fraction 1, no bubble; with dynamical
        orientation maximum angular speed 0 and rotation agility 1, not snapped), then its range (follow_range),
        lifetime (goto x), owner, the clock, MUSHROOM; filed. Freed (set_free), it grows at its turn in each tick and
        is removed there when its lifetime has passed. A second ball 0 in a world (AddBall's "not created" path) is
        refused with an error."""
        mask = mask.to(self.dev)
        if bool((mask & (self.ids != EMPTY_ID)).any()):
            raise ValueError("add_mushroom: a row is not empty")
        rng = f32(range_.to(device=self.dev, dtype=F64))
        life = life.to(device=self.dev, dtype=F64)
        orow = self.lookup(owner.to(device=self.dev, dtype=I64))
        ok = mask & (rng > 0.0) & (life > 0.0) & (orow >= 0)
        if not bool(ok.any()):
            return
        if bool((ok & (self.lookup(torch.zeros(self.n, dtype=I64, device=self.dev)) >= 0)).any()):
            raise ValueError("add_mushroom: ball 0 is in the world already (AddBall's not-created path)")
        if int(torch.bincount(self.world[ok]).max()) > 1:
            raise ValueError("add_mushroom: two mushrooms in one world")
        osafe = orow.clamp(min=0)
        where = lambda new, old: torch.where(ok.reshape(-1, *([1] * (old.dim() - 1))), new, old)
        zero3 = torch.zeros_like(self.pos)
        self.ids = torch.where(ok, 0, self.ids)
        self._index = None
        self.mass = torch.where(ok, MUSHROOM_MASS, self.mass)
        self.radius = torch.where(ok, 0.0, self.radius)
        self.max_vel = torch.where(ok, 0.0, self.max_vel)
        self.agility = torch.where(ok, 1.0, self.agility)
        self.speed_fraction = torch.where(ok, 1.0, self.speed_fraction)
        self.pos = where(self.pos[osafe], self.pos)
        self.old_pos = where(self.pos[osafe], self.old_pos)
        self.vel = where(zero3, self.vel)
        self.old_vel = where(zero3, self.old_vel)
        g = zero3.clone()
        g[:, 0] = life
        self.goto = where(g, self.goto)
        self.last_g = where(zero3, self.last_g)
        self.time_factor = torch.where(ok, self._decay(self.mass * self.agility), self.time_factor)
        self.free = torch.where(ok, False, self.free)
        self.massive = torch.where(ok, True, self.massive)
        self.interactive = torch.where(ok, False, self.interactive)
        self.moribund = torch.where(ok, False, self.moribund)
        self.junk = torch.where(ok, False, self.junk)
        self.sensor = torch.where(ok, False, self.sensor)
        self.cloaked = torch.where(ok, 0, self.cloaked)
        self.bubble = torch.where(ok, -1, self.bubble)
        self.formation_id = torch.where(ok, -1, self.formation_id.to(I64)).to(torch.int8)
        self.follow_range = torch.where(ok, rng, self.follow_range)
        self.owner_id = torch.where(ok, owner.to(device=self.dev, dtype=I64), self.owner_id)
        self.follow_id = torch.where(ok, 0, self.follow_id)
        self.follow_ptr = torch.where(ok, -1, self.follow_ptr)
        self.effect_stamp = torch.where(ok, self.clock, self.effect_stamp)
        self.last_collision = torch.where(ok, -1.0, self.last_collision)
        self.mode = torch.where(ok, MUSHROOM, self.mode)
        self._mush = self._mush | ok
        if self.dynamical_orientation:
            ident = torch.tensor([0.0, 0.0, 0.0, 1.0], dtype=F64, device=self.dev).expand(self.n, 4)
            self.rot = where(ident, self.rot)
            self.ang_vel = where(zero3, self.ang_vel)
            self.torque = where(zero3, self.torque)
            self.rot_roll = torch.where(ok, 0.0, self.rot_roll)
            self.max_ang_vel = torch.where(ok, 0.0, self.max_ang_vel)
            self.rot_agility = torch.where(ok, 1.0, self.rot_agility)
            self._orient_setup(ok, 1)              # the old rotation (at rest: nothing to snap)
        if self.collisions:
            self.last_c = where(zero3, self.last_c)
            self._refile(ok)
        self._state = None

    # ---- cloaking (VecCombatPark's API, src/exact_ballpark/vec_combat.py:497-519; full.py:1744-1780) ----

    def cloak(self, mask: torch.Tensor, mode: torch.Tensor, range_: torch.Tensor | None = None) -> None:
        """CloakBall for the balls in `mask`, a client's view (no proximity sensor: `range_` is the master's and
        changes nothing here): refused unless the mode, a signed char, is positive; the followers released
        (StopAllFollowers); cloaked with the mode's byte; non-massive."""
        mode = ((mode.to(device=self.dev, dtype=I64) + 128) & 0xFF) - 128
        ok = mask.to(self.dev) & (mode > 0)
        self._release_followers(ok)
        self.cloaked = torch.where(ok, mode & 0xFF, self.cloaked)
        self.massive = torch.where(ok, False, self.massive)

    def uncloak(self, mask: torch.Tensor) -> None:
        """UncloakBall for the balls in `mask`, a client's view: uncloaked; massive again unless in warp proper."""
        ok = mask.to(self.dev)
        self.cloaked = torch.where(ok, 0, self.cloaked)
        warping = (self.mode == WARP) & (self.effect_stamp >= 0)
        self.massive = torch.where(ok & ~warping, True, self.massive)

    # ---- formations (VecCombatPark's API, src/exact_ballpark/vec_combat.py:400-457; full.py:1859-1918) ----

    def load_formations(self, formations: list) -> None:
        """LoadFormations: the park's formations, each a list of slot offsets [x, y, z] (doubles). Without dynamical
        orientation a leader's rotation is its display orientation's: the park needs display_ypr=True."""
        f = [[[float(c) for c in v] for v in vectors] for vectors in formations]
        stride = max([len(v) for v in f] + [1])
        table = torch.zeros(max(1, len(f)), stride, 3, dtype=F64)
        for i, vectors in enumerate(f):
            if vectors:
                table[i, :len(vectors)] = torch.tensor(vectors, dtype=F64)
        lens = torch.tensor([len(v) for v in f] or [0], dtype=torch.int32)
        self._forms = {"table": table.to(self.dev).contiguous(), "len": lens.to(self.dev), "n": len(f),
                       "stride": stride, "lens": [len(v) for v in f]}
        self._state = None

    def _formation_len(self, fid: torch.Tensor) -> torch.Tensor:
        lens = torch.tensor(self._forms["lens"] or [0], dtype=I64, device=self.dev)
        return lens[fid.clamp(0, max(0, self._forms["n"] - 1))]

    def set_ball_formation(self, mask: torch.Tensor, formation: torch.Tensor) -> None:
        """SetBallFormation (a client's view) for the balls in `mask` (the formation a char): leaving a formation (-1)
        releases its followers (StopAllFollowers); a smaller one stops the FORMATION followers in the slots it lacks;
        a formation past the table is refused. A change from a formation gone past the table, or to a negative one, is
        Destiny's undefined behaviour (Ballpark.cpp:6696): refused with an error."""
        if self._forms is None:
            raise ValueError("set_ball_formation: load_formations first")
        fid = ((formation.to(device=self.dev, dtype=I64) + 128) & 0xFF) - 128
        F = self._forms["n"]
        ok = mask.to(self.dev) & ~(fid > F - 1)
        cur = self.formation_id.to(I64)
        inside = ok & (cur >= 0)
        leave = inside & (fid == -1)
        change = inside & (fid != -1) & (fid != cur)
        if bool((change & ((fid < 0) | (cur > F - 1))).any()):
            raise ValueError("SetBallFormation reads mFormations out of its range: Destiny's result is undefined")
        self._release_followers(leave)
        self._fslots.masked_fill_(leave, 0)             # leaving a formation clears every reservation
        new_len = self._formation_len(fid)
        shrink = change & (self._formation_len(cur) > new_len)
        safe = torch.where(self.follow_ptr >= 0, self.follow_ptr, 0)
        f = ((self.mode == FORMATION) & (self.follow_ptr >= 0) & shrink[safe]
             & (self.follow_id == self.ids[safe]) & (self.effect_stamp > new_len[safe] - 1))
        self._stop_rows(f)
        self.formation_id = torch.where(ok, fid, cur).to(torch.int8)

    @property
    def formation_slots(self) -> torch.Tensor:
        """[n] int64: each ball's reserved formation slots (Ball::mFormationSlots, a std::bitset<16>): set by
        reserve_formation_slot, cleared by free_formation_slot, by leaving its formation and, on the master, by the
        Stop(Ball*) of a FORMATION follower in that slot (its commands' and StopAllFollowers inside the tick)."""
        return self._fslots

    def _slot_leader(self, row: int, what: str) -> tuple:
        fid = int(self.formation_id[int(row)])
        if self._forms is None or fid == -1 or fid > self._forms["n"] - 1:
            return -1, 0
        if fid < 0:
            raise ValueError(f"{what} reads mFormations at a negative index: Destiny's result is undefined")
        return fid, self._forms["lens"][fid]

    def reserve_formation_slot(self, row: int) -> int:
        """ReserveFormationSlot (Ball.cpp:2275; full.py:1920-1936) of the ball in `row`: the first slot of its
        formation (of at most 16) not reserved, now reserved, or -1 (no formation, or more than 16 slots, or all
        taken). One ball on the host, as the reference's VecCombatPark does it (vec_combat.py:458-476)."""
        fid, n = self._slot_leader(row, "ReserveFormationSlot")
        if fid < 0 or n > FORMATION_MAX_SLOTS:
            return -1
        bits = int(self._fslots[int(row)])
        for i in range(n):
            if not bits >> i & 1:
                self._fslots[int(row)] = bits | (1 << i)
                return i
        return -1

    def free_formation_slot(self, row: int, slot: int) -> None:
        """FreeFormationSlot (Ball.cpp:2324; full.py:1938-1954) of the ball in `row`, with a size_t slot: nothing
        without a formation or for a slot past its end; a slot of 16 or more in a longer formation is Destiny's
        bitset<16> throwing: refused with an error."""
        fid, n = self._slot_leader(row, "FreeFormationSlot")
        if fid < 0:
            return
        s = int(slot) & 0xFFFFFFFFFFFFFFFF
        if s > ((n - 1) & 0xFFFFFFFFFFFFFFFF):
            return
        if s >= FORMATION_MAX_SLOTS:
            raise ValueError("FreeFormationSlot of a slot past 15: Destiny's bitset<16> throws")
        self._fslots[int(row)] = int(self._fslots[int(row)]) & ~(1 << s)

    def formation_follow(self, mask: torch.Tensor, leader: torch.Tensor, slot: torch.Tensor) -> None:
        """FormationFollow for the balls in `mask`: leader is each row's leader's Destiny id, slot a char; refused
        without the leader's formation or past its slots; stopped; then, unless the leader is in another bubble or
        cloaked, in the slot. A negative slot is Destiny's undefined behaviour (EvolveFormation reads before the
        formation, Ballpark.cpp:1044): refused with an error."""
        if self._forms is None:
            raise ValueError("formation_follow: load_formations first")
        if not self.dynamical_orientation and not self.display_ypr:
            raise ValueError("formations turn by the leader's rotation: without dynamical orientation that is the "
                             "display orientation (make the park with display_ypr=True)")
        if self.tiles is None:
            raise ValueError("formations need the tiled launch (each world's rows consecutive, in one workgroup)")
        slot = ((slot.to(device=self.dev, dtype=I64) + 128) & 0xFF) - 128
        dst = leader.to(device=self.dev, dtype=I64)
        lrow = self.lookup(dst)
        lsafe = lrow.clamp(min=0)
        fid = self.formation_id.to(I64)[lsafe]
        ok = (mask.to(self.dev) & (self.ids != dst) & (lrow >= 0) & (fid >= 0) & ~(fid > self._forms["n"] - 1)
              & ~(slot > self._formation_len(fid) - 1))
        if bool((ok & (slot < 0)).any()):
            raise ValueError("FormationFollow into a negative slot: EvolveFormation's result is undefined")
        if self.master and bool((ok & (slot >= FORMATION_MAX_SLOTS)).any()):
            raise ValueError("FormationFollow into a slot past 15 on the master: the follower's Stop would free it in "
                             "Destiny's bitset<16>, which throws")
        self._refuse_mushroom_targets(ok, lsafe)
        self._stop_rows(ok)
        ok = ok & (self.bubble == self.bubble[lsafe]) & (self.cloaked[lsafe] == 0)
        self.follow_id = torch.where(ok, dst, self.follow_id)
        self.follow_ptr = torch.where(ok, lrow, self.follow_ptr)
        self.effect_stamp = torch.where(ok, slot, self.effect_stamp)
        self.mode = torch.where(ok, FORMATION, self.mode)

    def warp_to(self, mask: torch.Tensor, point: torch.Tensor, min_range: torch.Tensor,
                warp_factor: torch.Tensor) -> None:
        x = torch.cat((point.to(device=self.dev, dtype=F64).reshape(self.n, 3),
                       min_range.to(device=self.dev, dtype=F64).reshape(self.n, 1)), dim=1)
        self._command("warp", mask, ints=(warp_factor,), x=x)

    def commands(self, op: torch.Tensor, point: torch.Tensor, target: torch.Tensor, range_: torch.Tensor,
                 min_range: torch.Tensor, warp_factor: torch.Tensor, rows: torch.Tensor | None = None) -> None:
        """At most one command per row in one launch (xb_modes_commands): op [n] uint8 (0 none, 3 stop, 1 goto_point,
        4 follow, 5 orbit, 6 warp_to) with point [n, 3], target [n] (a row), range_ [n], min_range [n] and
        warp_factor [n], each row's own arguments. `rows` (int32): only those rows (others' ops are not read). The
        result is that of the masked commands above applied op by op (stop, goto_point, follow, orbit, warp_to): a
        command writes only its own row and reads of its target only what no command writes."""
        n, d = self.n, self.dev
        args = [op.to(device=d, dtype=U8).contiguous(), point.to(device=d, dtype=F64).contiguous(),
                target.to(device=d, dtype=I64).contiguous(), range_.to(device=d, dtype=F64).contiguous(),
                min_range.to(device=d, dtype=F64).contiguous(), warp_factor.to(device=d, dtype=I64).contiguous()]
        if any(a.shape[0] != n for a in args) or args[1].shape != (n, 3):
            raise ValueError("commands: every argument has one entry (point: one [3] row) per park row")
        if rows is not None:
            rows = rows.to(device=d, dtype=torch.int32).contiguous()
        nr = n if rows is None else rows.numel()
        ptrs = [a.data_ptr() for a in args]
        s = self.state()
        if d.type == "cuda":
            with torch.cuda.device(d):
                rc = N.hip().xb_modes_commands(ctypes.byref(s), None if rows is None else rows.data_ptr(), nr, *ptrs,
                                               torch.cuda.current_stream(d).cuda_stream)
            N.check(rc, "xb_modes_commands")
        else:
            rc = N.cpu().xb_cpu_modes_commands(ctypes.byref(s), None if rows is None else rows.data_ptr(), nr, *ptrs)
            N.check(rc, "xb_cpu_modes_commands")

    def entity_warp_in(self, mask: torch.Tensor, point: torch.Tensor, warp_factor: torch.Tensor) -> None:
        self._command("warp_in", mask, ints=(warp_factor,), x=point, release=True)

    # ---- setters (vec_orient.py:403-466; SetBallVelocity without dynamical orientation, vec_modes.py:421-428) ----

    def set_mass(self, mask: torch.Tensor, mass: torch.Tensor) -> None:
        self._command("set_mass", mask, x=mass)

    def set_agility(self, mask: torch.Tensor, agility: torch.Tensor) -> None:
        self._command("set_agility", mask, x=agility)

    def _count_refiles(self, rec: torch.Tensor | None, mask: torch.Tensor) -> None:
        """`refiled`: the SetMaxSpeed and SetBallRadius calls that took their value, one a row."""
        if rec is None:
            return
        rows = torch.nonzero(mask.to(self.dev).reshape(self.n)).reshape(-1)
        self.refiled.index_add_(0, rows, ((rec[:, CMD_STATE_WORD] & CMD_ACCEPTED) != 0).to(I64))

    def set_max_velocity(self, mask: torch.Tensor, v: torch.Tensor) -> None:
        self._count_refiles(self._command("set_max_vel", mask, x=v), mask)

    def set_speed_fraction(self, mask: torch.Tensor, s: torch.Tensor) -> None:
        self._command("set_speed_fraction", mask, x=s)

    def set_radius(self, mask: torch.Tensor, r: torch.Tensor) -> None:
        self._count_refiles(self._command("set_radius", mask, x=r), mask)

    def set_velocity(self, mask: torch.Tensor, v: torch.Tensor) -> None:
        """SetBallVelocity: the new and old velocity; with dynamical orientation the rotation snapped to it, with the
        display orientation kept that snapped (full.py:1583-1593)."""
        self._command("set_velocity", mask, x=v)

    # ---- dynamical orientation's setters (vec_orient.py:438-464) ----

    def _orient_setup(self, mask: torch.Tensor | None, what: int) -> None:
        """xb_modes_orient_setup on the rows in mask (None: every row): 0 SnapOrientation, 1 a new ball's set-up."""
        if self.n == 0:
            return
        rows = None
        if mask is not None:
            rows = torch.nonzero(mask.to(self.dev).reshape(self.n)).reshape(-1).to(torch.int32)
            if rows.numel() == 0:
                return
        nr = self.n if rows is None else rows.numel()
        s = self.state()
        ptr = None if rows is None else rows.data_ptr()
        if self.dev.type == "cuda":
            with torch.cuda.device(self.dev):
                rc = N.hip().xb_modes_orient_setup(ctypes.byref(s), ptr, nr, int(what),
                                                   torch.cuda.current_stream(self.dev).cuda_stream)
            N.check(rc, "xb_modes_orient_setup")
        else:
            N.check(N.cpu().xb_cpu_modes_orient_setup(ctypes.byref(s), ptr, nr, int(what)),
                    "xb_cpu_modes_orient_setup")

    def set_max_angular_speed(self, mask: torch.Tensor, w: torch.Tensor) -> None:
        """SetMaxAngularSpeed: the float ignored if negative."""
        self._command("set_max_ang_speed", mask, x=w)

    def set_rot_agility(self, mask: torch.Tensor, a: torch.Tensor) -> None:
        """SetBallAngularAgility: ignored unless positive (a float)."""
        self._command("set_rot_agility", mask, x=a)

    def set_angular_velocity(self, mask: torch.Tensor, w: torch.Tensor) -> None:
        """SetBallAngularVelocity: the new and old angular velocity, as floats."""
        self._command("set_ang_vel", mask, x=w)

    def set_rotation_ypr(self, mask: torch.Tensor, ypr: torch.Tensor) -> None:
        """SetBallRotation(RotationQuaternion(yaw, pitch, roll)) with ypr [n, 3] (yaw, pitch, roll; each as a float)."""
        self._command("set_rotation_ypr", mask, x=ypr)

    def set_rotation(self, mask: torch.Tensor, q: torch.Tensor) -> None:
        """SetBallRotation: the float quaternion [n, 4] normalised when longer than 1e-20 (otherwise ignored), the
        rotation stopped."""
        self._command("set_rotation", mask, x=q)

    def set_free(self, mask: torch.Tensor, flag: torch.Tensor) -> None:
        """SetBallFree (vec_modes.py:438-451; full.py:1555-1572 without minis); an empty row is no ball: unchanged."""
        self._command("set_free", mask, flag=(flag.to(self.dev) != 0))

    def set_ball_rigid(self, mask: torch.Tensor) -> None:
        self._command("set_rigid", mask)

    def set_boid(self, mask: torch.Tensor, on: torch.Tensor) -> None:
        self._command("set_boid", mask, flag=(on.to(self.dev) != 0))

    def set_interactive(self, mask: torch.Tensor, flag: torch.Tensor) -> None:
        self._command("set_interactive", mask, flag=(flag.to(self.dev) != 0))

    def set_ball_troll(self, mask: torch.Tensor, delay: torch.Tensor) -> None:
        self._command("set_troll", mask, ints=(delay,))

    # ---- one tick (Ballpark::Evolve; vec_modes.py:943-960) ----

    def _tick(self, block: int | None = None, ticks: int = 1, active: torch.Tensor | None = None) -> None:
        """`ticks` ticks in one launch; `active` (a bool or uint8 device tensor of at least n rows, kept alive by the
        caller until the launch has run): OnTick's mask, the rows outside it skipped by the kernels."""
        s = self.state()
        if self.n == 0:
            return
        b = self.block if block is None else int(block)
        if ticks > 1 and self.tiles is None:
            for _ in range(ticks):
                self._tick(block, 1, active)
            return
        s.ticks = int(ticks)
        s.active = active.data_ptr() if active is not None else None
        try:
            if self.dev.type == "cuda":
                with torch.cuda.device(self.dev):
                    rc = N.hip().xb_modes_tick(ctypes.byref(s), torch.cuda.current_stream(self.dev).cuda_stream, b)
                N.check(rc, "xb_modes_tick")
            else:
                rc = N.cpu().xb_cpu_modes_tick(ctypes.byref(s), int(getattr(self, "cpu_threads", 0)))
                if rc:
                    raise RuntimeError(f"xb_cpu_modes_tick failed ({rc})")
        finally:
            s.active = None

    def _blocks(self) -> tuple:
        return ((self._f, self._q, self._b) + ((self._cf, self._cq, self._hits) if self.collisions else ())
                + ((self._y,) if self.display_ypr else ()))

    def _save(self) -> tuple:
        return tuple(b.clone() for b in self._blocks())

    def _put_back(self, saved: tuple, active: torch.Tensor) -> None:
        act = active.to(device=self.dev, dtype=torch.bool).reshape(self.n)
        sel = torch.zeros(self.stride, dtype=torch.bool, device=self.dev)
        sel[:self.n] = act
        for blk, old in zip(self._blocks(), saved):
            blk.copy_(torch.where(sel[None, :], blk, old))

    def evolve(self, lib=None, active: torch.Tensor | None = None) -> None:
        """One tick of every world, in place, no host synchronisation. `active` [n]: OnTick's per-world mask (a world's
        rows together): only those rows tick; every other row keeps every field bit for bit (the reference's
        masked_tick, vec_orient.py:217-254, which ticks every row and puts the others back; here the kernels skip the
        rows outside the mask: the tick reads only rows of each row's own world)."""
        if lib is not None:
            raise ValueError("lib: this park binds its own library")
        if active is not None:
            self._tick(active=active.to(device=self.dev, dtype=torch.bool).reshape(self.n).contiguous())
        else:
            self._tick()
        self.current_time += 1
        self._ticked = None if active is None else (self.current_time,
                                                    active.to(device=self.dev, dtype=torch.bool).reshape(self.n))

    def evolve_ticks(self, ticks: int) -> None:
        """`ticks` ticks of every world with no command between them: one launch when the park is tiled (each lane
        keeps its row in registers across the ticks), bit for bit the same as `ticks` calls of evolve()."""
        self._tick(ticks=int(ticks))
        self.current_time += int(ticks)

    def evolve_torch(self, active: torch.Tensor | None = None) -> None:
        """The reference's CPU tick name: here the same kernel source built for the host (a CPU park's evolve)."""
        if self.dev.type == "cuda":
            raise ValueError("evolve_torch is the CPU park's tick; a GPU park ticks with evolve()")
        self.evolve(active=active)

    # ---- OnTick (Ballpark::OnTick; full.py:2590-2620, vec_orient.py:834-879) ----

    def _world_times_disagree(self, sim: torch.Tensor) -> bool:
        """Whether two rows of one world hold different simulation times (OnTick runs whole worlds)."""
        _, wk = torch.unique(self.world, return_inverse=True)
        lo = torch.full((int(wk.max()) + 1,), torch.iinfo(torch.int64).max, dtype=I64, device=self.dev)
        hi = torch.full_like(lo, torch.iinfo(torch.int64).min)
        lo.scatter_reduce_(0, wk, sim, reduce="amin")
        hi.scatter_reduce_(0, wk, sim, reduce="amax")
        return bool((lo != hi).any())

    def on_tick(self, sim_time, tick=None, after=None) -> list:
        """Ballpark::OnTick for every world at its own simulation time (`sim_time` [n] in 100 ns units, a world's rows
        sharing it): BringOutTheDeads first, in every world; then per world the intervals it is due, an interval being
        the tick interval x 10,000: none when less than one has passed since the world's `time` (its time kept); one
        on its first call (its time then `sim_time`); on a master's park (`master=True`) one when more than
        MAX_CATCH_UP_TICKS are due (the backlog dropped); otherwise every whole interval, the remainder carried in its
        time. Round j ticks the worlds due more than j intervals together (`tick(rows)`, by default
        `evolve(active=rows)`, `tick(None)` when every row is due); `after(rows)` follows each round. Returns each
        round's row mask."""
        n, d = self.n, self.dev
        if getattr(self, "time", None) is None:          # Be::Time mTime and mFirstTime, as a new ballpark holds them
            self.time = torch.zeros(n, dtype=I64, device=d)
            self.first_time = torch.ones(n, dtype=torch.bool, device=d)
        sim = torch.as_tensor(sim_time, dtype=I64).to(d).reshape(n)
        if n and self._world_times_disagree(sim):
            raise ValueError("on_tick: a world's rows hold different simulation times")
        self.bring_out_the_deads()
        step = int(self.tick_interval) * 10000
        behind = sim - self.time
        due = behind >= step
        whole = torch.where(due, torch.div(behind, step, rounding_mode="floor"), 0)
        fresh = due & self.first_time                   # a first call evolves once
        resync = (whole > MAX_CATCH_UP_TICKS) if self.master else torch.zeros_like(due)   # a master drops a backlog
        rounds = torch.where(fresh | resync, 1, whole)
        carry = torch.where(fresh | resync, 0, behind - whole * step)
        total = int(rounds.max()) if n else 0
        tick = tick or (lambda rows: self.evolve(active=rows))
        masks = []
        for j in range(total):
            rows = rounds > j
            tick(None if bool(rows.all()) else rows)
            if after is not None:
                after(rows)
            masks.append(rows)
        self.first_time = self.first_time & ~fresh
        self.time = torch.where(due, sim - carry, self.time)
        return masks


class ModesGraph:
    """A VecModesPark's tick captured once as a HIP graph on the park's own state blocks, so each replay is one tick
    (the reference's ModesGraph, src/exact_ballpark/vec_modes_cuda.py:371-432); with `ticks` ticks a launch and
    `launches` launches a replay, one replay runs ticks x launches ticks back to back. Commands and setters that write
    into the blocks need no re-binding between replays; a call that replaces a buffer the launches point at (the first
    set_harmonic, load_formations, load_statics, record_paths, a regrouping, new tiles) makes tick() raise: capture
    again. With `masked=True` the capture holds OnTick's per-world mask as a device buffer the kernels read (the rows
    outside it skipped); `tick(active)` writes the mask and replays."""

    def __init__(self, park: VecModesPark, block: int | None = None, lib=None, masked: bool = False,
                 ticks: int = 1, launches: int = 1):
        if park.dev.type != "cuda":
            raise ValueError("a HIP graph needs a park on a GPU")
        if lib is not None:
            raise ValueError("lib: the park binds its own library")
        self.park, self.block = park, park.block if block is None else int(block)
        self.masked = masked
        self.ticks, self.launches = int(ticks), int(launches)
        if masked and (self.ticks != 1 or self.launches != 1):
            raise ValueError("OnTick's mask applies to one tick at a time")
        if self.launches < 1 or self.ticks < 1:
            raise ValueError("ticks and launches are at least 1")
        self.mask = torch.ones(park.stride, dtype=torch.bool, device=park.dev)
        act = self.mask if masked else None
        before = park._save()                            # warm-up: the module loaded before the capture,
        park._tick(self.block, self.ticks, act)          # the park's state put back after it
        for b, old in zip(park._blocks(), before):
            b.copy_(old)
        torch.cuda.synchronize(park.dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for _ in range(self.launches):
                park._tick(self.block, self.ticks, act)
        self.captured = self._pointers()

    def tick(self, active: torch.Tensor | None = None) -> None:
        if active is not None and not self.masked:
            raise ValueError("a graph made without OnTick's mask (masked=True) ticks every row")
        if self.masked:
            if active is None:
                self.mask.fill_(True)
            else:
                self.mask[:self.park.n].copy_(active.to(device=self.park.dev, dtype=torch.bool).reshape(-1))
        if self._pointers() != self.captured:
            raise RuntimeError("a block, table or buffer of the park was replaced since the capture (a first "
                               "set_harmonic, load_formations, load_statics, record_paths, a regrouping or new tiles): "
                               "capture the graph again")
        self.graph.replay()
        self.park.current_time += self.ticks * self.launches
        self.park._ticked = None if not self.masked or active is None else (
            self.park.current_time, self.mask[:self.park.n].clone())

    def _pointers(self) -> tuple:
        """Every pointer of the park's state the capture froze into its launches."""
        s = self.park.state()
        return tuple(getattr(s, name) for name, typ in N.ModesState._fields_ if typ is ctypes.c_void_p) + (
            s.n_tiles, s.hits_cap, s.n_form)
