"""Scene files in the reference's .psy format, parsed by libpsy_gpu.so
(psy_parse_scene; parser/parser.cpp, parser/data_tree.cpp).

    scene = psyfile.load("frame.psy")          # the first Scene node
    g = gpu.GpuScene(ctx, scene)               # the description it holds
    W, H, spp, seed = (scene.settings[k] for k in ("width", "height", "spp", "seed"))
"""
import ctypes as C

from . import abi
from .gpu import check, lib


def _text(src):
    return src.encode("utf-8") if isinstance(src, str) else bytes(src)


def scene_count(text):
    """Number of Scene nodes (frames) in .psy text."""
    b = _text(text)
    n = C.c_int32()
    check(lib().psy_parse_scene_count(b, len(b), C.byref(n)), "psy_parse_scene_count")
    return n.value


def dump_tree(text):
    """The DataTree of .psy text as the reference's print_tree prints it."""
    b = _text(text)
    need = C.c_int64()
    check(lib().psy_parse_dump_tree(b, len(b), None, 0, C.byref(need)), "psy_parse_dump_tree")
    buf = C.create_string_buffer(need.value)
    check(lib().psy_parse_dump_tree(b, len(b), buf, need.value, C.byref(need)), "psy_parse_dump_tree")
    return buf.value.decode("utf-8")


class ParsedScene:
    """One parsed Scene node: a scene description for GpuScene (or any
    consumer of psy_scene_desc), its render settings and the parser's
    warnings."""

    def __init__(self, text, scene_index=0):
        b = _text(text)
        h = C.c_void_p()
        check(lib().psy_parse_scene(b, len(b), int(scene_index), C.byref(h)), "psy_parse_scene")
        self.handle = h
        s = abi.RenderSettings()
        check(lib().psy_parsed_scene_settings(h, C.byref(s)), "psy_parsed_scene_settings")
        self.settings = {"width": s.width, "height": s.height, "spp": s.spp, "seed": s.seed,
                         "dice_rate": s.dice_rate, "name": s.name.decode("utf-8", "replace"),
                         "output_path": s.output_path.decode("utf-8", "replace")}
        self.warnings = lib().psy_parsed_scene_warnings(h).decode("utf-8", "replace")

    def desc(self):
        """The psy_scene_desc (a ctypes structure owned by this object)."""
        return lib().psy_parsed_scene_desc(self.handle).contents

    def ptr(self):
        if not self.handle:
            raise ValueError("parsed scene is closed")
        return lib().psy_parsed_scene_desc(self.handle)

    def to_desc(self):
        return self

    def close(self):
        if self.handle:
            lib().psy_parsed_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load(path, scene_index=0):
    """Parse Scene node `scene_index` of a .psy file."""
    with open(path, "rb") as f:
        return ParsedScene(f.read(), scene_index)
