"""The reference's .ren script front-end (xyz/core/shell.cpp:195-497) over the
C ABI: the same commands, defaults and output names, with the path tracer of
xyz_render as the renderer.

Commands (one per line; blank lines, "rem ..." lines and =begin/=end blocks
are skipped):
    set rep N        integrator: 0 path tracing (PT/), 1 importance-driven path
                     tracing (IDPT/; the importon map of 100,000 light paths is
                     built once before the first pass, as ImportanceDrivenPathTracer
                     does), 8 bidirectional path tracing (BPT/; own and light-
                     tracing contributions in one film); the reference's other
                     integrators are not part of this library
    set rt <time>    rendering time, e.g. "1 h 30 min", "10 s" (units d, h, m, s
                     and their long forms, as stream2time)
    set it <time>    interval between film snapshots
    set eye x y z    camera position (metres)
    set at x y z     gazing point
    set dim W H      image size (default 720 x 480)
    load NAME        loads <mqo_dir>/NAME.mqo and builds its kd-tree
    save NAME        writes <ren_dir>/NAME.ren reproducing the current settings
    ren              renders until the rendering time is spent, writing
                     "<out_dir>/<scene> WxH <date>/<PT|IDPT|BPT>/<hh>h<mm>m<ss>s <spp>spp
                     <queries>.pfm" at every interval and at the end
    quit / exit / abort
The reference renders one spp pass over the image at a time and checks the
clock between passes; here a pass is xyz_render over a batch of samples whose
size follows the measured rate, so that snapshots land near the intervals.
Camera: the reference's thin lens (f = 35 mm, F1.4) at eye / at; path depth
cap 16 vertices (max_depth 15), Russian roulette at every vertex, threshold 0.2
(xyz/core/config.hpp:30,60-61)."""
import datetime
import os
import time

import numpy as np

from . import film as film_io
from . import mqo as mqo_mod

MAX_DEPTH = 15
RR_THRESHOLD = 0.2
N_EMIT = 100000  # CONFIG_EMITED_SIZE (xyz/core/config.hpp:57)
# set rep N -> output folder (xyz/core/shell.cpp:391-437)
INTEGRATORS = {0: "PT", 1: "IDPT", 8: "BPT"}

_UNITS = {}
for _names, _secs in ((("d", "day", "days"), 86400), (("h", "hr", "hour", "hours"), 3600),
                      (("m", "min", "minute", "minutes"), 60), (("s", "sec", "second", "seconds"), 1)):
    for _n in _names:
        _UNITS[_n] = _secs


def stream2time(tokens):
    """Seconds of "<int> <unit> [<int> <unit> ...]" (shell.cpp stream2time);
    an unknown unit adds nothing, a token that is not an integer ends the list."""
    total = 0
    i = 0
    while i < len(tokens):
        try:
            val = int(tokens[i])
        except ValueError:
            break
        unit = tokens[i + 1] if i + 1 < len(tokens) else ""
        total += val * _UNITS.get(unit, 0)
        i += 2
    return total


def time2string(t):
    """shell.cpp time2string: "1 day 2 hours 3 minutes 4 seconds"; zero is " 0 s"."""
    parts = []
    for n, one, many in ((t // 86400, " day", " days"), ((t // 3600) % 24, " hour", " hours"),
                         ((t // 60) % 60, " minute", " minutes"), (t % 60, " second", " seconds")):
        if n > 0:
            parts.append(f"{n}{many if n > 1 else one}")
    return " ".join(parts) if parts else " 0 s"


class Shell:
    """Interpreter state: scene, camera, dimension, times, integrator."""

    def __init__(self, mqo_dir=".", out_dir="out", ren_dir="ren", seed=0, log=print,
                 clock=time.monotonic):
        self.mqo_dir, self.out_dir, self.ren_dir = mqo_dir, out_dir, ren_dir
        self.seed = seed
        self.log = log
        self.clock = clock
        self.scene_name = ""
        self.mqo = None
        self.scene = None
        self.eye = [0.0, 1.0, 4.0]
        self.at = [0.0, 1.0, 0.0]
        self.width, self.height = 720, 480
        self.rendering_time = 0
        self.interval_time = 0
        self.algorithm = 0
        self.outputs = []

    # ---------------------------------------------------------------- commands
    def interpret(self, line):
        tok = line.split()
        if not tok:
            return "continue"
        cmd = tok[0]
        if cmd in ("quit", "abort"):
            return "quit"
        if cmd == "exit":
            return "exit"
        if cmd == "set" and len(tok) >= 2:
            what, args = tok[1], tok[2:]
            if what == "rep":
                self.algorithm = int(args[0])
            elif what == "rt":
                self.rendering_time = stream2time(args)
                self.interval_time = min(self.interval_time, self.rendering_time)
            elif what == "it":
                self.interval_time = stream2time(args)
                self.rendering_time = max(self.rendering_time, self.interval_time)
            elif what == "eye":
                self.eye = [float(x) for x in args[:3]]
            elif what == "at":
                self.at = [float(x) for x in args[:3]]
            elif what == "dim":
                self.width = max(1, int(args[0]))
                self.height = max(1, int(args[1]))
            return "continue"
        if cmd == "load" and len(tok) >= 2:
            self.load(tok[1])
            return "continue"
        if cmd == "save" and len(tok) >= 2:
            self.save(tok[1])
            return "continue"
        if cmd == "ren":
            if self.rendering_time <= 0:
                return "continue"
            self.render()
            return "quit"
        return "undefined"

    def load(self, name):
        path = os.path.join(self.mqo_dir, name + ".mqo")
        if not os.path.exists(path):
            self.log(f"load: {path} not found")
            return
        self.scene_name = name
        self.mqo = mqo_mod.load(path)
        if self.mqo.info["has_camera"]:
            self.eye, self.at = list(self.mqo.info["eye"]), list(self.mqo.info["at"])
        self.scene = self.mqo.scene()
        self.log(f"load scene ({path}): {self.mqo.n_tris} triangles, "
                 f"kd build {self.scene.info()['build_seconds']:.3f}s")

    def save(self, name):
        os.makedirs(self.ren_dir, exist_ok=True)
        with open(os.path.join(self.ren_dir, name + ".ren"), "w") as f:
            f.write(f"load {self.scene_name}\n\n")
            f.write("set eye " + " ".join(f"{x:g}" for x in self.eye) + "\n")
            f.write("set at  " + " ".join(f"{x:g}" for x in self.at) + "\n\n")
            f.write(f"set dim {self.width} {self.height}\n")
            f.write(f"set it  {time2string(self.interval_time)}\n")
            f.write(f"set rt  {time2string(self.rendering_time)}\n\n")
            f.write(f"set rep {self.algorithm}\nren\n")

    def render(self):
        if self.scene is None:
            raise RuntimeError("ren: no scene loaded")
        if self.algorithm not in INTEGRATORS:
            raise NotImplementedError(
                f"ren: integrator {self.algorithm} is not provided (0 = path tracing, "
                "1 = importance-driven path tracing, 8 = bidirectional path tracing)")
        from . import hip
        sub = INTEGRATORS[self.algorithm]
        stamp = datetime.datetime.now().strftime("%Y-%m-%dT%H.%M.%S")
        outdir = os.path.join(self.out_dir, f"{self.scene_name} {self.width:04d}x{self.height:04d} "
                                            f"{stamp} ", sub)
        os.makedirs(outdir, exist_ok=True)
        W, H = self.width, self.height
        cam = [*self.eye, *self.at, 0.0, 1.0, 0.0, 0.0, mqo_mod.FOCAL_LENGTH, mqo_mod.F_NUMBER]
        film = hip.DeviceArray.zeros((H, W, 3), np.float32)
        spp, queries = 0, 0
        batch = 1
        interval = max(1, self.interval_time or self.rendering_time)
        next_save = interval
        imap = None
        if sub == "IDPT":
            imap = self.scene.importon_map(N_EMIT, max_depth=MAX_DEPTH, seed=self.seed)
        t0 = self.clock()
        while True:
            t1 = self.clock()
            if sub == "PT":
                st = self.scene.render(cam, film, batch, MAX_DEPTH, rr_start=0,
                                       rr_threshold=RR_THRESHOLD, seed=self.seed,
                                       sample_offset=spp)
            elif sub == "IDPT":
                st = self.scene.render_idpt(imap, cam, film, batch, MAX_DEPTH, rr_start=0,
                                            rr_threshold=RR_THRESHOLD, seed=self.seed,
                                            sample_offset=spp)
            else:
                st = self.scene.render_bdpt(cam, film, batch, MAX_DEPTH, seed=self.seed,
                                            sample_offset=spp)
            hip.synchronize()
            spp += batch
            queries += int(st["queries"])
            now = self.clock()
            elapsed = now - t0
            done = elapsed >= self.rendering_time
            if done or elapsed >= next_save:
                sec = int(elapsed + 0.5)
                name = (f"{sec // 3600:02d}h{(sec // 60) % 60:02d}m{sec % 60:02d}s "
                        f"{spp:06d}spp {queries:012d}.pfm")
                path = os.path.join(outdir, name)
                film_io.save_pfm(path, film.numpy(), spp)
                self.outputs.append(path)
                self.log(f"  {name}")
                next_save += interval
            if done:
                break
            # size the next pass to about a tenth of an interval
            rate = batch / max(now - t1, 1e-6)
            batch = int(max(1, min(4096, rate * min(interval, self.rendering_time) / 10)))
        return self.outputs

    # ---------------------------------------------------------------- scripts
    def run(self, lines):
        """Interprets lines until quit/exit or the end; returns the last status."""
        it = iter(lines)
        status = "continue"
        for raw in it:
            line = raw.strip()
            if line == "=begin":
                for raw2 in it:
                    if raw2.strip() == "=end":
                        break
                continue
            if not line or line.startswith("rem "):
                continue
            status = self.interpret(line)
            if status == "undefined":
                self.log(f"undefined: {line}")
            if status in ("quit", "exit"):
                break
        return status


def run_script(path, **kw):
    sh = Shell(**kw)
    with open(path) as f:
        sh.run(f.read().splitlines())
    return sh
