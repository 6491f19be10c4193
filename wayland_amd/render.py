"""Headless frame loop and outputs: the reference's Renderer
(src/Frontend/Core/Renderer.cpp) with the window taken out.

    python -m wayland_amd.render -s scene/cornell-box/cbox.pbrt -fr 64 -o out -r 512 512

* Command line as RendererSetting(argc, argv) (Renderer.cpp:157-191):
  -fr frame limit, -t time limit in seconds, -s scene, -o output folder
  (which also turns output on), -p program, -r width height.  Defaults as
  SetDefault (Renderer.cpp:193-209): no limits, PathTracing, 1000 x 1000,
  no output.  An unknown -p name is ignored, as the reference ignores it.
* Run (Renderer.cpp:28-58): Setup sets frame 0 and a zeroed float4 radiance
  buffer (PathTracingProgramManager.cpp:11-38); each iteration launches one
  progressive frame (the running mean and ACES u8 frame buffer of
  PathTracing.cu:140-159, through wl_pt_launch_frame), downloads the frame
  buffer (DeviceManager.cpp:243-248) and advances the frame id (Update,
  PathTracingProgramManager.cpp:40-47); the loop ends when the time limit
  (checked in whole milliseconds) or the frame limit is reached.  Without a
  window to close, a run with neither limit is refused.
* WriteOutput (Renderer.cpp:60-106): "<out>/<scene>-<program>-<time>-[<ctime>]"
  with ':' -> '_' in the first two places, a .txt of run statistics in the
  reference's format, and a .png of the frame buffer flipped vertically
  (row 0 of the buffer is the bottom of the image).
Only the PathTracing program is on this backend; the other program names of
Renderer.h:32-36 raise.
"""
from __future__ import annotations

import os
import struct
import sys
import time
import zlib
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

PROGRAM_NAME = ("AO", "Color", "Depth", "Mesh", "Normal", "Simple", "Texture", "WireFrame",
                "PathTracing", "RandomWalk", "BDPT")


@dataclass
class RendererSetting:
    frame_limit: int = -1
    time_limit: float = -1.0
    scene_path: Optional[str] = None
    output_path: str = "."
    program: str = "PathTracing"
    resolution: tuple = (1000, 1000)
    write_output: bool = False

    @classmethod
    def from_args(cls, argv: Sequence[str]) -> "RendererSetting":
        s = cls()
        i = 0
        argv = list(argv)
        while i < len(argv):
            a = argv[i]
            if a == "-fr":
                i += 1
                s.frame_limit = int(argv[i])
            elif a == "-t":
                i += 1
                s.time_limit = float(argv[i])
            elif a == "-s":
                i += 1
                s.scene_path = argv[i]
            elif a == "-o":
                i += 1
                s.output_path = argv[i]
                s.write_output = True
            elif a == "-p":
                i += 1
                if argv[i] in PROGRAM_NAME:
                    s.program = argv[i]
            elif a == "-r":
                s.resolution = (int(argv[i + 1]), int(argv[i + 2]))
                i += 2
            i += 1
        return s

    @property
    def scene_name(self) -> str:
        """File name without folder and extension (SetSceneName)."""
        return os.path.splitext(os.path.basename(self.scene_path or ""))[0]


def render_time_label(seconds: float) -> str:
    """std::to_string of the float, cut one digit after the point, plus 's'."""
    txt = "%f" % float(np.float32(seconds))
    return txt[:txt.index(".") + 2] + "s"


def time_stamp(now: Optional[float] = None) -> str:
    """ctime() without the newline, the first two ':' replaced by '_', in brackets."""
    s = time.ctime(time.time() if now is None else now)
    s = s.replace(":", "_", 2)
    return f"[{s}]"


def stats_text(scene_name, program, resolution, render_time, frame_count) -> str:
    """The .txt the reference writes; floats as an ostream prints them (%g)."""
    rt = np.float32(render_time)
    ft = np.float32(rt / np.float32(frame_count))
    fps = np.float32(np.float32(frame_count) / rt)
    return ("{\n"
            f"\tScene: {scene_name}\n"
            f"\tProgram: {program}\n"
            f"\tResolution: {resolution[0]} {resolution[1]}\n"
            f"\tRenderTime: {float(rt):g}\n"
            f"\tFrameCount: {frame_count}\n"
            f"\tFrameTime: {float(ft):g}\n"
            f"\tFramesPerSecond: {float(fps):g}\n"
            "}\n")


def png_bytes(rgba: np.ndarray) -> bytes:
    """8-bit RGBA PNG of an (H, W, 4) array, first row at the top."""
    rgba = np.ascontiguousarray(rgba, np.uint8)
    h, w, c = rgba.shape
    if c != 4:
        raise ValueError("expected 4 channels")

    def chunk(tag, data):
        body = tag + data
        return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body))

    raw = np.concatenate([np.zeros((h, 1), np.uint8), rgba.reshape(h, w * 4)], 1)
    return (b"\x89PNG\r\n\x1a\n"
            + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw.tobytes(), 6))
            + chunk(b"IEND", b""))


def save_image_cpu(path: str, width: int, height: int, rgba: np.ndarray) -> None:
    """examples/Common/Common.cpp:23-41 SaveImageCPU: the RGBA8 buffer as a
    JPEG of quality 90, rows in buffer order (stbi_write_jpg drops alpha).
    Encoded with Pillow, so the bytes differ from stb's encoder."""
    from PIL import Image
    img = np.asarray(rgba, np.uint8).reshape(height, width, 4)
    Image.fromarray(np.ascontiguousarray(img[..., :3]), "RGB").save(path, "JPEG", quality=90)


def flip_frame_buffer(frame_buffer: np.ndarray, width: int, height: int) -> np.ndarray:
    """The whole buffer reversed, then each row reversed back: a vertical
    flip (Renderer.cpp:96-102)."""
    return np.asarray(frame_buffer, np.uint8).reshape(height, width, 4)[::-1]


class Renderer:
    """Scene + PathTracing program + device buffers, as Renderer's
    constructor sets them up (Renderer.cpp:20-26)."""

    def __init__(self, setting: RendererSetting, ctx=None, max_depth: int = 25,
                 rr_depth: int = 5):
        if setting.program != "PathTracing":
            raise NotImplementedError(f"program '{setting.program}' is not on this backend")
        if not setting.scene_path:
            raise ValueError("no scene (-s)")
        import torch

        import wayland_amd as wl
        from . import pbrt
        self.setting = setting
        self.width, self.height = setting.resolution
        self.loaded = pbrt.load_pbrt(setting.scene_path, self.width, self.height)
        self.own_ctx = ctx is None
        self.ctx = ctx or wl.Context(0)
        self.accel = wl.Accel(self.ctx, self.loaded.scene)
        self.pt = wl.PathTracer(self.ctx, self.accel, self.loaded.scene)
        self.cfg = wl.pt_config(self.width, self.height, max_depth=max_depth,
                                rr_depth=rr_depth)
        npix = self.width * self.height
        self.radiance = torch.zeros((npix, 4), dtype=torch.float32, device="cuda")
        # two colour buffers and two pinned host buffers: frame k's download
        # (on its own stream) overlaps the render of frame k + 1
        self.color = [torch.zeros((npix, 4), dtype=torch.uint8, device="cuda") for _ in range(2)]
        self.host = [torch.empty((npix, 4), dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        self.copy_stream = torch.cuda.Stream()
        self.copied = [None, None]  # event: the download from colour buffer k is done
        self.frame_buffer = np.zeros((npix, 4), np.uint8)
        self.frame_id = 0

    def run(self):
        """Frames until a limit; returns (frames, seconds)."""
        s = self.setting
        if s.frame_limit <= 0 and s.time_limit <= 0:
            raise ValueError("a headless run needs a frame limit (-fr) or a time limit (-t)")
        import torch
        self.frame_id = 0
        t0 = time.perf_counter_ns()
        frames = 0
        main = torch.cuda.current_stream()
        last = None  # (slot, event) of the newest download
        while True:
            slot = self.frame_id & 1
            if self.copied[slot] is not None:  # that buffer's previous download
                main.wait_event(self.copied[slot])
            self.pt.launch_frame(self.cfg, self.frame_id, self.radiance, self.color[slot])
            rendered = torch.cuda.Event()
            rendered.record(main)
            self.copy_stream.wait_event(rendered)
            with torch.cuda.stream(self.copy_stream):
                self.host[slot].copy_(self.color[slot], non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.copy_stream)
            self.copied[slot] = done
            if last is not None:  # the previous frame is on the host by now
                last[1].synchronize()
                self.frame_buffer = self.host[last[0]].numpy()
            last = (slot, done)
            self.frame_id += 1
            frames += 1
            dur_ns = time.perf_counter_ns() - t0
            out_time = s.time_limit > 0 and s.time_limit * 1000 < dur_ns // 1_000_000
            out_frame = s.frame_limit > 0 and frames >= s.frame_limit
            if out_time or out_frame:
                break
        last[1].synchronize()  # the last frame, kept past later runs
        self.frame_buffer = self.host[last[0]].numpy().copy()
        dur_ns = time.perf_counter_ns() - t0  # up to the last frame on the host
        seconds = float(np.float32(dur_ns // 1000) / np.float32(1000) / np.float32(1000))
        if s.write_output:
            self.write_output(seconds, frames)
        return frames, seconds

    def write_output(self, render_time: float, frame_count: int, now: Optional[float] = None):
        s = self.setting
        base = os.path.join(s.output_path, f"{s.scene_name}-{s.program}-"
                            f"{render_time_label(render_time)}-{time_stamp(now)}")
        with open(base + ".txt", "w") as f:
            f.write(stats_text(s.scene_name, s.program, s.resolution, render_time, frame_count))
        with open(base + ".png", "wb") as f:
            f.write(png_bytes(flip_frame_buffer(self.frame_buffer, self.width, self.height)))
        return base

    def close(self):
        self.pt.close()
        self.accel.close()
        if self.own_ctx:
            self.ctx.close()


def main(argv=None) -> int:
    setting = RendererSetting.from_args(sys.argv[1:] if argv is None else argv)
    r = Renderer(setting)
    try:
        frames, seconds = r.run()
        print(f"{frames} frames in {seconds:.3f} s ({frames / max(seconds, 1e-9):.1f} fps)")
    finally:
        r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
