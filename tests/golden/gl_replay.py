"""Golden-vector generator, part 2 of 2 (development container only).

Replays a WebGL2 command trace written by ``ref_trace.js`` on SwiftShader,
the CPU implementation of OpenGL ES 3.0 that ships with the ``kaleido``
package, through EGL pbuffer contexts driven by ctypes.  The SwiftShader
libraries are run, never inspected.  ``__dump`` records in the trace read
textures back with ``glReadPixels`` and are written as ``<tag>.<name>.npy``
(float32, shape (h, w, 4)).

Usage: python gl_replay.py <trace dir> [--swiftshader DIR]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

SS_DEFAULT = "/usr/local/lib/python3.10/dist-packages/kaleido/executable/bin/swiftshader"

EGL_DEFAULT_DISPLAY = 0
EGL_SURFACE_TYPE, EGL_PBUFFER_BIT = 0x3033, 0x0001
EGL_RENDERABLE_TYPE, EGL_OPENGL_ES3_BIT = 0x3040, 0x0040
EGL_RED_SIZE, EGL_GREEN_SIZE, EGL_BLUE_SIZE, EGL_ALPHA_SIZE = 0x3024, 0x3023, 0x3022, 0x3021
EGL_NONE, EGL_WIDTH, EGL_HEIGHT = 0x3038, 0x3057, 0x3056
EGL_CONTEXT_CLIENT_VERSION, EGL_OPENGL_ES_API = 0x3098, 0x30A0
GL_COMPILE_STATUS, GL_LINK_STATUS, GL_INFO_LOG_LENGTH = 0x8B81, 0x8B82, 0x8B84
GL_FRAMEBUFFER, GL_READ_FRAMEBUFFER, GL_COLOR_ATTACHMENT0 = 0x8D40, 0x8CA8, 0x8CE0
GL_TEXTURE_2D, GL_RGBA, GL_FLOAT, GL_FRAMEBUFFER_COMPLETE = 0x0DE1, 0x1908, 0x1406, 0x8CD5
GL_FRAMEBUFFER_BINDING = 0x8CA6


# SwiftShader 4.1 mishandles sampler-typed function arguments in divergent
# control flow: a fragment whose 2x2 quad also holds fragments on the other
# side of a branch reads the sampler argument of the other side's call (or
# none, giving zero).  GLSL ES 3.00 gives those calls the same meaning as a
# texelFetch at the call site, so the three monopole helpers of the physics
# shaders (GLSL_MASS, src/glsl-sim.js:57-66) are rewritten to take the texel
# instead of the sampler before compiling -- a rewrite that keeps every
# operation of the reference's arithmetic.
SAMPLER_ARG_REWRITES = [
    ("vec3 massAt(vec3 p, sampler2D m) { vec4 mono = texelFetch(m, ivec2(0), 0);", "vec3 massAt(vec3 p, vec4 mono) {"),
    ("float potAt(vec3 p, sampler2D m) { vec4 mono = texelFetch(m, ivec2(0), 0);", "float potAt(vec3 p, vec4 mono) {"),
    ("vec3 crossAt(sampler2D other, sampler2D mine) {\n  vec4 mo = texelFetch(other, ivec2(0), 0), mm = texelFetch(mine, ivec2(0), 0);",
     "vec3 crossAt(vec4 mo, vec4 mm) {"),
    ("massAt(p, uMonoB)", "massAt(p, texelFetch(uMonoB, ivec2(0), 0))"),
    ("massAt(p, uMono)", "massAt(p, texelFetch(uMono, ivec2(0), 0))"),
    ("potAt(p, uMonoB)", "potAt(p, texelFetch(uMonoB, ivec2(0), 0))"),
    ("potAt(p, uMono)", "potAt(p, texelFetch(uMono, ivec2(0), 0))"),
    ("crossAt(uMonoB, uMono)", "crossAt(texelFetch(uMonoB, ivec2(0), 0), texelFetch(uMono, ivec2(0), 0))"),
    ("crossAt(uMono, uMonoB)", "crossAt(texelFetch(uMono, ivec2(0), 0), texelFetch(uMonoB, ivec2(0), 0))"),
]


def rewrite_sampler_args(src):
    if "sampler2D m)" not in src and "crossAt(sampler2D" not in src:
        return src
    for a, b in SAMPLER_ARG_REWRITES:
        src = src.replace(a, b)
    if "sampler2D m)" in src or "crossAt(sampler2D" in src:
        raise RuntimeError("sampler-argument rewrite incomplete")
    return src


class Replayer:
    def __init__(self, ss_dir):
        self.egl = C.CDLL(os.path.join(ss_dir, "libEGL.so"))
        self.gl = C.CDLL(os.path.join(ss_dir, "libGLESv2.so"))
        e = self.egl
        e.eglGetDisplay.restype = C.c_void_p
        e.eglGetDisplay.argtypes = [C.c_void_p]
        e.eglInitialize.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        e.eglChooseConfig.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_int)]
        e.eglCreatePbufferSurface.restype = C.c_void_p
        e.eglCreatePbufferSurface.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        e.eglCreateContext.restype = C.c_void_p
        e.eglCreateContext.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        e.eglMakeCurrent.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        dpy = e.eglGetDisplay(None)
        major, minor = C.c_int(), C.c_int()
        if not e.eglInitialize(dpy, C.byref(major), C.byref(minor)):
            raise RuntimeError("eglInitialize failed")
        attrs = (C.c_int * 15)(EGL_SURFACE_TYPE, EGL_PBUFFER_BIT, EGL_RENDERABLE_TYPE, EGL_OPENGL_ES3_BIT,
                               EGL_RED_SIZE, 8, EGL_GREEN_SIZE, 8, EGL_BLUE_SIZE, 8, EGL_ALPHA_SIZE, 8, EGL_NONE, 0, 0)
        cfg = C.c_void_p()
        n = C.c_int()
        if not e.eglChooseConfig(dpy, attrs, C.byref(cfg), 1, C.byref(n)) or n.value < 1:
            raise RuntimeError("eglChooseConfig failed")
        surf = e.eglCreatePbufferSurface(dpy, cfg, (C.c_int * 5)(EGL_WIDTH, 16, EGL_HEIGHT, 16, EGL_NONE))
        e.eglBindAPI(EGL_OPENGL_ES_API)
        ctx = e.eglCreateContext(dpy, cfg, None, (C.c_int * 3)(EGL_CONTEXT_CLIENT_VERSION, 3, EGL_NONE))
        if not ctx or not e.eglMakeCurrent(dpy, surf, surf, ctx):
            raise RuntimeError("eglCreateContext/MakeCurrent failed")
        self.gl.glGetString.restype = C.c_char_p
        self.version = self.gl.glGetString(0x1F02).decode()
        self.renderer = self.gl.glGetString(0x1F01).decode()
        self.ids = {}          # trace handle -> GL name
        self.kind = {}
        self.uloc = {}         # (prog handle, name) -> location
        self.cur_prog = None
        self.cur_fb = 0
        self.verbose = bool(os.environ.get("GL_REPLAY_VERBOSE"))
        self.unit = 0
        self.bound = {}        # texture unit -> trace handle
        self.clear_color = (0.0, 0.0, 0.0, 0.0)
        self.prog_fs = {}      # program handle -> fragment shader source
        self.cell_draws = 0
        self.marks = []        # (tag, wall time) at each __mark

    # -- helpers --
    def gen(self, fn, h, kind):
        v = C.c_uint()
        fn(1, C.byref(v))
        self.ids[h] = v.value
        self.kind[h] = kind

    def real(self, a):
        if a is None:
            return 0
        return self.ids[a["h"]]

    def data_ptr(self, a, blobs):
        if a is None:
            return None
        return C.c_void_p(blobs.ctypes.data + a["blob"])

    def loc(self, a):
        if a is None:
            return -1
        key = (a["prog"], a["loc"])
        if key not in self.uloc:
            self.uloc[key] = self.gl.glGetUniformLocation(self.ids[a["prog"]], a["loc"].encode())
        return self.uloc[key]

    def check_shader(self, sid, src):
        ok = C.c_int()
        self.gl.glGetShaderiv(sid, GL_COMPILE_STATUS, C.byref(ok))
        if not ok.value:
            buf = C.create_string_buffer(8192)
            self.gl.glGetShaderInfoLog(sid, 8192, None, buf)
            raise RuntimeError("shader compile failed: " + buf.value.decode() + "\n" + src[:400])

    def dump(self, rec, out_dir):
        g = self.gl
        fb = C.c_uint()
        g.glGenFramebuffers(1, C.byref(fb))
        g.glBindFramebuffer(GL_FRAMEBUFFER, fb.value)
        for d in rec["dump"]:
            g.glFramebufferTexture2D(GL_FRAMEBUFFER, GL_COLOR_ATTACHMENT0, GL_TEXTURE_2D, self.ids[d["tex"]], 0)
            st = g.glCheckFramebufferStatus(GL_FRAMEBUFFER)
            if st != GL_FRAMEBUFFER_COMPLETE:
                raise RuntimeError("dump framebuffer incomplete for %s: %x" % (d["name"], st))
            g.glReadBuffer(GL_COLOR_ATTACHMENT0)
            arr = np.zeros((d["h"], d["w"], 4), np.float32)
            g.glReadPixels(0, 0, d["w"], d["h"], GL_RGBA, GL_FLOAT, arr.ctypes.data_as(C.c_void_p))
            np.save(os.path.join(out_dir, "%s.%s.npy" % (rec["tag"], d["name"])), arr)
        g.glBindFramebuffer(GL_FRAMEBUFFER, self.cur_fb)
        g.glDeleteFramebuffers(1, C.byref(fb))

    def is_cell_pass(self):
        fs = self.prog_fs.get(self.cur_prog, "")
        return "uniform sampler2D uFarT0;" in fs and "me.y >= 512" in fs

    def read_tex(self, h, w, hh):
        g = self.gl
        fb = C.c_uint()
        g.glGenFramebuffers(1, C.byref(fb))
        g.glBindFramebuffer(GL_FRAMEBUFFER, fb.value)
        g.glFramebufferTexture2D(GL_FRAMEBUFFER, GL_COLOR_ATTACHMENT0, GL_TEXTURE_2D, self.ids[h], 0)
        arr = np.zeros((hh, w, 4), np.float32)
        g.glReadPixels(0, 0, w, hh, GL_RGBA, GL_FLOAT, arr.ctypes.data_as(C.c_void_p))
        g.glBindFramebuffer(GL_FRAMEBUFFER, self.cur_fb)
        g.glDeleteFramebuffers(1, C.byref(fb))
        return arr

    def sparse_cell_draw(self, a):
        """The per-cell gravity pass is a full-screen draw over the 512x513
        fine-mesh atlas whose fragments return at once for empty cells;
        SwiftShader still walks their loops, which makes the one draw take
        tens of minutes.  Each fragment depends only on its own coordinate,
        so the same draw restricted by scissor boxes to the occupied cells
        (and the scratch row, which the shader copies) writes the same values
        there; every other texel of the target is what the shader writes for
        an empty cell, zero."""
        g = self.gl
        fine = self.read_tex(self.bound[0], 512, 513)
        occ = fine[:512, :, 3] > 0.0
        g.glClearColor(C.c_float(0), C.c_float(0), C.c_float(0), C.c_float(0))
        g.glClear(0x4000)
        cc = self.clear_color
        g.glClearColor(C.c_float(cc[0]), C.c_float(cc[1]), C.c_float(cc[2]), C.c_float(cc[3]))
        GL_SCISSOR_TEST = 0x0C11
        g.glEnable(GL_SCISSOR_TEST)
        for y in range(512):
            row = occ[y]
            if not row.any():
                continue
            xs = np.flatnonzero(row)
            start = xs[0]
            prev = xs[0]
            for x in list(xs[1:]) + [None]:
                if x is not None and x == prev + 1:
                    prev = x
                    continue
                g.glScissor(int(start), y, int(prev - start + 1), 1)
                t0 = time.time()
                g.glDrawArrays(a[0], a[1], a[2])
                self.cell_draws += 1
                if self.verbose and self.cell_draws <= 5:
                    g.glFinish()
                    print("  cell span y=%d x=%d len=%d %.3fs (occupied %d)" % (y, start, prev - start + 1, time.time() - t0, occ.sum()), flush=True)
                if x is not None:
                    start = prev = x
        g.glScissor(0, 512, 512, 1)
        g.glDrawArrays(a[0], a[1], a[2])
        g.glDisable(GL_SCISSOR_TEST)

    def run(self, trace_dir):
        tr = json.load(open(os.path.join(trace_dir, "trace.json")))
        blobs = np.fromfile(os.path.join(trace_dir, "blobs.bin"), np.uint8)
        shaders = {int(k): v for k, v in tr["shaders"].items()}
        g = self.gl
        f = C.c_float
        timing = {}
        for rec in tr["trace"]:
            op, a = rec["op"], rec["args"]
            t0 = time.time()
            if op == "__dump":
                g.glFinish()
                self.dump(rec, trace_dir)
                continue
            if op == "__mark":   # every earlier command done, then the time noted
                g.glFinish()
                self.marks.append((rec["tag"], time.time()))
                continue
            if op == "createShader":
                self.ids[rec["ret"]] = g.glCreateShader(a[0])
            elif op == "shaderSource":
                src = rewrite_sampler_args(shaders[a[0]["h"]]).encode()
                arr = (C.c_char_p * 1)(src)
                g.glShaderSource(self.real(a[0]), 1, arr, None)
            elif op == "compileShader":
                g.glCompileShader(self.real(a[0]))
                self.check_shader(self.real(a[0]), shaders[a[0]["h"]])
            elif op == "createProgram":
                self.ids[rec["ret"]] = g.glCreateProgram()
            elif op == "attachShader":
                g.glAttachShader(self.real(a[0]), self.real(a[1]))
                src = shaders[a[1]["h"]]
                if "gl_FragCoord" in src or "out vec4 o" in src:
                    self.prog_fs[a[0]["h"]] = src
            elif op == "linkProgram":
                g.glLinkProgram(self.real(a[0]))
                ok = C.c_int()
                g.glGetProgramiv(self.real(a[0]), GL_LINK_STATUS, C.byref(ok))
                if not ok.value:
                    raise RuntimeError("link failed")
            elif op in ("getShaderParameter", "getProgramParameter", "getActiveUniform", "getUniformLocation",
                        "getShaderInfoLog", "getProgramInfoLog", "getExtension", "getParameter",
                        "checkFramebufferStatus", "fenceSync", "deleteSync", "clientWaitSync", "flush",
                        "getBufferSubData", "readPixels", "readBuffer"):
                pass
            elif op == "createTexture":
                self.gen(g.glGenTextures, rec["ret"], "tex")
            elif op == "createFramebuffer":
                self.gen(g.glGenFramebuffers, rec["ret"], "fb")
            elif op == "createRenderbuffer":
                self.gen(g.glGenRenderbuffers, rec["ret"], "rb")
            elif op == "createBuffer":
                self.gen(g.glGenBuffers, rec["ret"], "buf")
            elif op == "createVertexArray":
                self.gen(g.glGenVertexArrays, rec["ret"], "vao")
            elif op == "bindVertexArray":
                g.glBindVertexArray(self.real(a[0]))
            elif op == "bindTexture":
                self.bound[self.unit] = a[1]["h"] if a[1] else None
                g.glBindTexture(a[0], self.real(a[1]))
            elif op == "texParameteri":
                g.glTexParameteri(a[0], a[1], a[2])
            elif op == "texImage2D":
                g.glTexImage2D(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], self.data_ptr(a[8], blobs))
            elif op == "texSubImage2D":
                g.glTexSubImage2D(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], self.data_ptr(a[8], blobs))
            elif op == "bindFramebuffer":
                self.cur_fb = self.real(a[1])
                g.glBindFramebuffer(a[0], self.cur_fb)
            elif op == "framebufferTexture2D":
                g.glFramebufferTexture2D(a[0], a[1], a[2], self.real(a[3]), a[4])
            elif op == "framebufferRenderbuffer":
                g.glFramebufferRenderbuffer(a[0], a[1], a[2], self.real(a[3]))
            elif op == "drawBuffers":
                lst = a[0]
                g.glDrawBuffers(len(lst), (C.c_uint * len(lst))(*lst))
            elif op == "bindRenderbuffer":
                g.glBindRenderbuffer(a[0], self.real(a[1]))
            elif op == "renderbufferStorage":
                g.glRenderbufferStorage(a[0], a[1], a[2], a[3])
            elif op == "bindBuffer":
                g.glBindBuffer(a[0], self.real(a[1]))
            elif op == "bufferData":
                g.glBufferData(a[0], C.c_ssize_t(int(a[1])), None, a[2])
            elif op == "deleteTexture":
                g.glDeleteTextures(1, C.byref(C.c_uint(self.real(a[0]))))
            elif op == "deleteFramebuffer":
                g.glDeleteFramebuffers(1, C.byref(C.c_uint(self.real(a[0]))))
            elif op == "deleteRenderbuffer":
                g.glDeleteRenderbuffers(1, C.byref(C.c_uint(self.real(a[0]))))
            elif op == "useProgram":
                self.cur_prog = a[0]["h"]
                g.glUseProgram(self.real(a[0]))
            elif op == "uniform1i":
                g.glUniform1i(self.loc(a[0]), int(a[1]))
            elif op == "uniform2i":
                g.glUniform2i(self.loc(a[0]), int(a[1]), int(a[2]))
            elif op == "uniform1f":
                g.glUniform1f(self.loc(a[0]), f(a[1]))
            elif op == "uniform2f":
                g.glUniform2f(self.loc(a[0]), f(a[1]), f(a[2]))
            elif op == "uniform3f":
                g.glUniform3f(self.loc(a[0]), f(a[1]), f(a[2]), f(a[3]))
            elif op == "uniform3fv":
                v = [float(x) for x in a[1]]
                g.glUniform3fv(self.loc(a[0]), 1, (C.c_float * 3)(*v))
            elif op == "uniformMatrix3fv":
                v = [float(x) for x in a[2]]
                g.glUniformMatrix3fv(self.loc(a[0]), 1, 1 if a[1] else 0, (C.c_float * 9)(*v))
            elif op == "viewport":
                g.glViewport(*[int(x) for x in a])
            elif op == "enable":
                g.glEnable(a[0])
            elif op == "disable":
                g.glDisable(a[0])
            elif op == "depthFunc":
                g.glDepthFunc(a[0])
            elif op == "depthMask":
                g.glDepthMask(1 if a[0] else 0)
            elif op == "clearColor":
                self.clear_color = tuple(a)
                g.glClearColor(f(a[0]), f(a[1]), f(a[2]), f(a[3]))
            elif op == "clear":
                g.glClear(a[0])
            elif op == "colorMask":
                g.glColorMask(*[1 if x else 0 for x in a])
            elif op == "activeTexture":
                self.unit = a[0] - 0x84C0
                g.glActiveTexture(a[0])
            elif op == "blendFunc":
                g.glBlendFunc(a[0], a[1])
            elif op == "drawArrays" and self.is_cell_pass():
                self.sparse_cell_draw(a)
            elif op == "drawArrays":
                g.glDrawArrays(a[0], a[1], a[2])
                if self.verbose:
                    g.glFinish()
                    print("draw prog=%s mode=%d count=%d %.3fs" % (self.cur_prog, a[0], a[2], time.time() - t0), flush=True)
            else:
                raise RuntimeError("unhandled GL op " + op)
            err = g.glGetError()
            if err:
                raise RuntimeError("GL error 0x%x after %s" % (err, op))
            timing[op] = timing.get(op, 0.0) + time.time() - t0
        g.glFinish()
        return timing


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace_dir")
    ap.add_argument("--swiftshader", default=os.environ.get("SWIFTSHADER_DIR", SS_DEFAULT))
    args = ap.parse_args()
    r = Replayer(args.swiftshader)
    t0 = time.time()
    timing = r.run(args.trace_dir)
    info = {"renderer": r.renderer, "version": r.version, "wall_s": time.time() - t0,
            "per_op_s": {k: round(v, 4) for k, v in sorted(timing.items(), key=lambda kv: -kv[1])[:8]},
            "marks": [[tag, t - t0] for tag, t in r.marks]}
    json.dump(info, open(os.path.join(args.trace_dir, "replay.json"), "w"), indent=1)
    print(json.dumps(info))


if __name__ == "__main__":
    sys.exit(main())
