// Golden-vector generator, part 1 of 2 (runs in the development container
// only, never on the GPU host, and never imported by the product).
//
// The reference is a browser program: JavaScript host code that drives GLSL
// passes through WebGL2. This script evaluates the reference's own source
// files (read as text from REF_DIR, default /root/reference/src) in a node
// `vm` context with a *recording* WebGL2 object: every GL call the reference
// makes is written to a command trace, typed-array payloads to a blob file.
// gl_replay.py then replays that trace on SwiftShader (a CPU implementation
// of OpenGL ES 3.0) and writes the textures the trace asks to dump.
//
// The job file (JSON) lists engine operations to perform, in order:
//   {"op":"build", "opt":{...}}                 CC.nbody.build(opt)
//   {"op":"set_state", "dir":"...", "name":"s"} upload pos/vel/q/mat from .f32 files
//   {"op":"set_sim", "fields":{...}}            assign sim fields (t, bulk0, cen0, boxes...)
//   {"op":"set_pp", "list":"file.i32"}          the loose list (as setPP does)
//   {"op":"step", "settle":b, "kick":b}         CC.nbody.step(settle, kick)
//   {"op":"pass_sim", "settle":b,"kick":b,"gdt":x} CC.nbody.passes.sim
//   {"op":"diag"}                               CC.nbody.passes.diag (boxes, seats, mesh, pp, DIAG)
//   {"op":"rigid", "dtF":x}                     CC.nbody.rigidStep
//   {"op":"dump", "what":[...], "tag":"..."}    trace command: read textures back
//   {"op":"mark", "tag":"..."}                  trace command: finish the GPU work and note the time
// Host-side values (sim scalars, the pair table, uniforms) are written to
// <out>/host.json directly from the evaluated reference.
'use strict';
const fs = require('fs');
const path = require('path');
const vm = require('vm');

const REF = process.env.REF_DIR || '/root/reference/src';
const jobFile = process.argv[2];
const outDir = process.argv[3];
if (!jobFile || !outDir) { console.error('usage: node ref_trace.js job.json outdir'); process.exit(2); }
fs.mkdirSync(outDir, { recursive: true });
const job = JSON.parse(fs.readFileSync(jobFile, 'utf8'));

// ---- WebGL2 enum values (identical to the OpenGL ES 3.0 ones) ----
const ENUM = {
  TEXTURE_2D: 0x0DE1, TEXTURE_MIN_FILTER: 0x2801, TEXTURE_MAG_FILTER: 0x2800, NEAREST: 0x2600,
  TEXTURE_WRAP_S: 0x2802, TEXTURE_WRAP_T: 0x2803, CLAMP_TO_EDGE: 0x812F,
  RGBA: 0x1908, RED: 0x1903, RG: 0x8227, FLOAT: 0x1406, HALF_FLOAT: 0x140B,
  RGBA32F: 0x8814, R32F: 0x822E, RG32F: 0x8230, RGBA16F: 0x881A,
  FRAMEBUFFER: 0x8D40, COLOR_ATTACHMENT0: 0x8CE0, COLOR_ATTACHMENT1: 0x8CE1, COLOR_ATTACHMENT2: 0x8CE2,
  COLOR_ATTACHMENT3: 0x8CE3, DEPTH_ATTACHMENT: 0x8D00, RENDERBUFFER: 0x8D41, DEPTH_COMPONENT24: 0x81A6,
  FRAMEBUFFER_COMPLETE: 0x8CD5, DEPTH_TEST: 0x0B71, LESS: 0x0201, BLEND: 0x0BE2, ONE: 1,
  COLOR_BUFFER_BIT: 0x4000, DEPTH_BUFFER_BIT: 0x0100, POINTS: 0, TRIANGLES: 4,
  VERTEX_SHADER: 0x8B31, FRAGMENT_SHADER: 0x8B30, COMPILE_STATUS: 0x8B81, LINK_STATUS: 0x8B82,
  ACTIVE_UNIFORMS: 0x8B86, PIXEL_PACK_BUFFER: 0x88EB, STREAM_READ: 0x88E1,
  SYNC_GPU_COMMANDS_COMPLETE: 0x9117, ALREADY_SIGNALED: 0x911A, CONDITION_SATISFIED: 0x911C,
  ALIASED_POINT_SIZE_RANGE: 0x846D
};
for (let i = 0; i < 32; i++) ENUM['TEXTURE' + i] = 0x84C0 + i;

// ---- the recorder ----
const trace = [];
const blobs = [];
let blobBytes = 0;
let nextId = 1;
const shaderSrc = new Map();     // shader handle id -> source
const progShaders = new Map();   // program id -> [shader ids]
function handle(kind) { return { __h: nextId++, kind: kind }; }
function enc(a) {
  if (a === null || a === undefined) return null;
  if (typeof a === 'number' || typeof a === 'string' || typeof a === 'boolean') return a;
  if (a.__h) return { h: a.__h };
  if (a.__loc) return { loc: a.__loc, prog: a.__prog };
  if (ArrayBuffer.isView(a)) {
    const buf = Buffer.from(a.buffer, a.byteOffset, a.byteLength);
    const cn = a.constructor && a.constructor.name;   // typed arrays made inside the vm context are of another realm
    const type = cn === 'Float32Array' ? 'f32' : cn === 'Int32Array' ? 'i32' : cn === 'Uint8Array' ? 'u8' : 'raw';
    blobs.push(Buffer.from(buf));
    const rec = { blob: blobBytes, bytes: buf.length, type: type };
    blobBytes += buf.length;
    return rec;
  }
  if (Array.isArray(a)) return a.map(enc);
  throw new Error('cannot encode GL argument ' + a);
}
function uniformsOf(src) {
  const out = [];
  const re = /uniform\s+\w+\s+(\w+)\s*;/g;
  let m;
  while ((m = re.exec(src))) out.push(m[1]);
  return out;
}
const special = {
  createShader: () => handle('shader'), createProgram: () => handle('program'), createTexture: () => handle('texture'),
  createFramebuffer: () => handle('framebuffer'), createRenderbuffer: () => handle('renderbuffer'),
  createBuffer: () => handle('buffer'), createVertexArray: () => handle('vao'), createQuery: () => handle('query'),
  fenceSync: () => handle('sync'),
  getShaderParameter: () => true, getShaderInfoLog: () => '', getProgramInfoLog: () => '',
  checkFramebufferStatus: () => ENUM.FRAMEBUFFER_COMPLETE,
  clientWaitSync: () => ENUM.ALREADY_SIGNALED,
  getExtension: () => ({}),
  getParameter: (p) => (p === ENUM.ALIASED_POINT_SIZE_RANGE ? [1, 8192] : 0),
  getProgramParameter: (p, what) => {
    if (what === ENUM.ACTIVE_UNIFORMS) {
      const names = new Set();
      for (const s of progShaders.get(p.__h) || []) for (const n of uniformsOf(shaderSrc.get(s))) names.add(n);
      p.__uniforms = Array.from(names);
      return p.__uniforms.length;
    }
    return true;
  },
  getActiveUniform: (p, i) => ({ name: p.__uniforms[i] }),
  getUniformLocation: (p, name) => ({ __loc: name, __prog: p.__h })
};
// Per-box force injection (see install_injector): when the reference binds a
// mesh box's force framebuffer, texture unit 10 is re-bound to that box's
// injected force mesh, and the active unit put back.
let activeUnit = ENUM.TEXTURE0;
let injectFor = null;   // (framebuffer handle id) -> injected texture, or null
const gl = new Proxy({}, {
  get(target, prop) {
    if (typeof prop !== 'string') return undefined;
    if (prop in ENUM) return ENUM[prop];
    return function () {
      const args = Array.prototype.slice.call(arguments);
      if (prop === 'shaderSource') shaderSrc.set(args[0].__h, args[1]);
      if (prop === 'attachShader') { const l = progShaders.get(args[0].__h) || []; l.push(args[1].__h); progShaders.set(args[0].__h, l); }
      if (prop === 'activeTexture') activeUnit = args[0];
      let ret;
      if (special[prop]) ret = special[prop].apply(null, args);
      const rec = { op: prop, args: args.map(enc) };
      if (ret && ret.__h) rec.ret = ret.__h;
      trace.push(rec);
      if (prop === 'bindFramebuffer' && injectFor && args[1] && injectFor.has(args[1].__h)) {
        trace.push({ op: 'activeTexture', args: [ENUM.TEXTURE10] });
        trace.push({ op: 'bindTexture', args: [ENUM.TEXTURE_2D, enc(injectFor.get(args[1].__h))] });
        trace.push({ op: 'activeTexture', args: [activeUnit] });
      }
      return ret;
    };
  }
});

// ---- evaluate the reference in a browser-like context ----
const ctx = vm.createContext({ console: console, screen: { width: 1920, height: 1080 },
                               document: { getElementById: () => null } });
vm.runInContext('var window = this;', ctx);
for (const f of ['shell.js', 'math.js', 'bodies.js', 'glsl-common.js', 'glsl-sim.js', 'gl.js', 'nbody.js'])
  vm.runInContext(fs.readFileSync(path.join(REF, f), 'utf8'), ctx, { filename: f });
const CC = ctx.CC;
CC.gl.init({ gl: gl, floatBlend: true });
CC.nbody.init({ gl: gl, floatBlend: true });
const NB = CC.nbody;

// ---- the crater page (chicxulub.html), on the same recording GL ----
// The crater view is page code, not an engine module: its constants, shaders
// and functions live inside the page's script. This takes the pieces the
// crater's physics needs out of the page's text -- the constants and
// COUPLE_FS (from `var L_UNIT` to the globe section), and latticeFill,
// buildCrater, setStaticUniforms, passSlots, passSim, step, settleSome,
// approachStep and couple by name -- and evaluates them in one function
// scope beside stubs for the page's view, DOM and camera.
const HTML = process.env.REF_HTML || path.join(REF, '..', 'chicxulub.html');
let crater = null;
function craterPage() {
  const text = fs.readFileSync(HTML, 'utf8');
  const between = (a, b) => {
    const i = text.indexOf(a), j = text.indexOf(b, i);
    if (i < 0 || j < 0) throw new Error('crater page section not found: ' + a);
    return text.slice(i, j);
  };
  const fn = (name) => {
    const start = text.indexOf('function ' + name + '(');
    if (start < 0) throw new Error('crater page function not found: ' + name);
    let depth = 0;
    for (let i = text.indexOf('{', start); i < text.length; i++) {
      if (text[i] === '{') depth++;
      else if (text[i] === '}') { depth--; if (depth === 0) return text.slice(start, i + 1); }
    }
    throw new Error('unbalanced ' + name);
  };
  const frames = /var APPROACH_FRAMES = [^;]+;/.exec(text)[0];
  const src = [
    '(function (gl, hooks) {',
    "'use strict';",
    'var $ = function () { return { classList: { remove: function () {} }, hidden: false }; };',
    'var GLH = CC.gl, program = GLH.program, floatTex = GLH.floatTex;',
    'var simTarget = GLH.simTarget, freeTarget = GLH.freeTarget, makeGrid = GLH.makeGrid, freeGrid = GLH.freeGrid;',
    between('var L_UNIT', '// ---------- the first day'),
    'var simProg = program(QUAD_VS, SIM_FS), slotProg = program(SLOT_VS, SLOT_FS);',
    'var rigidProg = program(QUAD_VS, RIGID_FS), coupleProg = program(QUAD_VS, COUPLE_FS);',
    'var VIEW = { setPalette: function () {}, setMaxPoint: function () {}, setRadPow: function () {}, setLook: function () {}, clampDist: function () {} };',
    'var cam = {}, run = {}, MATH = CC.math, BODIES = CC.bodies, mulberry = MATH.mulberry, ball = BODIES.ball;',
    "var grid = null, sim = null, report = null, framedAt = 0, gen = 0, stepAcc = 0, lastReport = '';",
    fn('latticeFill'), fn('buildCrater'), fn('setStaticUniforms'), fn('passSlots'), fn('passSim'), fn('step'),
    fn('settleSome'), frames, fn('approachStep'), fn('couple'),
    'var coupleOf = couple;',
    'couple = function () { if (hooks.beforeCouple) hooks.beforeCouple(); coupleOf(); };',
    'return { S: S, build: buildCrater, settleSome: settleSome, step: step, approachStep: approachStep,',
    '         get sim() { return sim; }, get grid() { return grid; },',
    '         consts: { HH_C1: HH_C1, HH_MU: HH_MU, HH_P: HH_P, Z_MODEL: Z_MODEL, E_KG: E_KG, CP_ROCK: CP_ROCK,',
    '                   G_EARTH: G_EARTH, T_UNIT: T_UNIT, L_UNIT: L_UNIT, MAT_MELT: MAT_MELT, MAT_FROZEN: MAT_FROZEN,',
    '                   APPROACH_FRAMES: APPROACH_FRAMES } };',
    '})'
  ].join('\n');
  const hooks = {};
  const page = vm.runInContext(src, ctx, { filename: 'chicxulub.html#crater' })(gl, hooks);
  page.hooks = hooks;
  return page;
}

function texOf(name) {
  if (crater) {
    const s = crater.sim, g = crater.grid;
    switch (name) {
      case 'pos': return { tex: s.src.pos, w: s.W, h: s.H };
      case 'vel': return { tex: s.src.vel, w: s.W, h: s.H };
      case 'aux': return { tex: s.src.aux, w: s.W, h: s.H };
      case 'q': return { tex: s.src.q, w: s.W, h: s.H };
      case 'seatA': return { tex: g.slot[0], w: g.AW, h: g.AH };
      case 'seatB': return { tex: g.slot[1], w: g.AW, h: g.AH };
      default: throw new Error('unknown crater texture ' + name);
    }
  }
  const sim = NB.sim, pm = NB.pm, g = NB.grid, pp = NB.pp;
  const m = /^(fine|coarse|mono|farA|farT0|farT1|force)([01])$/.exec(name);
  if (m) {
    const dims = { fine: [512, 513], coarse: [64, 64], mono: [1, 1], farA: [64, 64], farT0: [64, 64], farT1: [64, 64], force: [512, 513] }[m[1]];
    return { tex: pm[+m[2]][m[1]], w: dims[0], h: dims[1] };
  }
  switch (name) {
    case 'pos': return { tex: sim.src.pos, w: sim.W, h: sim.H };
    case 'vel': return { tex: sim.src.vel, w: sim.W, h: sim.H };
    case 'aux': return { tex: sim.src.aux, w: sim.W, h: sim.H };
    case 'q': return { tex: sim.src.q, w: sim.W, h: sim.H };
    case 'diag': return { tex: sim.diag, w: sim.W, h: sim.H };
    case 'seatA': return { tex: g.slot[0], w: g.AW, h: g.AH };
    case 'seatB': return { tex: g.slot[1], w: g.AW, h: g.AH };
    case 'ppList': return { tex: pp.list, w: 128, h: 64 };
    case 'ppBox': return { tex: pp.box, w: 128, h: 64 };
    case 'ppForce': return { tex: pp.force, w: 128, h: 256 };
    default: throw new Error('unknown texture ' + name);
  }
}
function readF32(dir, name) {
  const b = fs.readFileSync(path.join(dir, name));
  return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4).slice();
}
function upload(tex, w, h, data, fmt) {
  gl.bindTexture(gl.TEXTURE_2D, tex);
  gl.texSubImage2D(gl.TEXTURE_2D, 0, 0, 0, w, h, fmt, gl.FLOAT, data);
}

const host = { ops: [] };
let injTex = null;
for (const op of job.ops) {
  const sim = NB.sim;
  switch (op.op) {
    case 'build': {
      NB.build(op.opt);
      const s = NB.sim;
      const keep = ['N', 'W', 'H', 'n1', 'n2', 'm', 'a', 'cell', 'dt', 'k', 'c', 'creep', 'R1', 'R2', 'Re1', 'Re2', 'M1', 'M2', 'M',
                    'vc', 'vesc', 'spin0', 'spin1', 'cen0', 'cen1', 'bulk0', 'bulk1', 't', 'settleLeft', 'damp', 'tK', 'kappa', 'rel'];
      const rec = {};
      for (const k of keep) rec[k] = s[k];
      rec.G = NB.grid.G;
      host.ops.push({ op: 'build', sim: rec });
      trace.push({ op: '__dump', args: [], dump: [texOf('pos'), texOf('vel')].map((t, i) => ({ name: ['pos', 'vel'][i], tex: t.tex.__h, w: t.w, h: t.h })), tag: op.tag || 'build' });
      fs.writeFileSync(path.join(outDir, (op.tag || 'build') + '.mat.f32'), Buffer.from(s.matR.buffer));
      fs.writeFileSync(path.join(outDir, (op.tag || 'build') + '.home.f32'), Buffer.from(s.homeArr.buffer));
      break;
    }
    case 'set_state': {
      const n = op.name;
      upload(sim.src.pos, sim.W, sim.H, readF32(op.dir, n + '.pos.f32'), gl.RGBA);
      upload(sim.src.vel, sim.W, sim.H, readF32(op.dir, n + '.vel.f32'), gl.RGBA);
      upload(sim.src.q, sim.W, sim.H, readF32(op.dir, n + '.q.f32'), gl.RED);
      const mat = readF32(op.dir, n + '.mat.f32');
      sim.matR = mat;
      upload(sim.mat, sim.W, sim.H, mat, gl.RED);
      break;
    }
    case 'set_sim': {
      for (const k of Object.keys(op.fields)) sim[k] = op.fields[k];
      break;
    }
    case 'set_mono': {   // prime a box's mono texel, as placeBoxes does
      upload(NB.pm[op.box].mono, 1, 1, new Float32Array(op.value), gl.RGBA);
      break;
    }
    case 'set_pp': {
      const b = fs.readFileSync(op.list);
      const list = new Int32Array(b.buffer, b.byteOffset, b.byteLength / 4);
      const pp = NB.pp;
      pp.slotArr.fill(-1);
      for (let k = 0; k < list.length; k++) pp.slotArr[list[k]] = k;
      pp.indexArr.fill(-1);
      for (let k = 0; k < list.length; k++) pp.indexArr[k] = list[k];
      upload(pp.slot, sim.W, sim.H, pp.slotArr, gl.RED);
      upload(pp.index, 128, 64, pp.indexArr, gl.RED);
      pp.count = list.length; pp.loose = list.length;
      break;
    }
    case 'step': NB.step(!!op.settle, !!op.kick); if (op.advance) sim.t += sim.dt; break;
    case 'pass_slots': NB.passes.slots(); break;
    case 'pass_sim': NB.passes.sim(!!op.settle, !!op.kick, op.gdt); break;
    case 'diag': NB.passes.diag(); break;
    case 'rigid': NB.rigidStep(op.dtF); break;
    case 'settle': NB.settleSome(op.n); break;
    case 'record_sim': {
      const rec = {};
      for (const k of op.fields) rec[k] = JSON.parse(JSON.stringify(sim[k]));
      host.ops.push({ op: 'record_sim', tag: op.tag, sim: rec });
      break;
    }
    case 'install_injector': {
      // The per-cell gravity pass (CELL_FS) is replaced, through the
      // reference's own setProg hook, by a copy of a texture on unit 10 that
      // holds a force mesh computed elsewhere: SwiftShader cannot run CELL_FS
      // in useful time. Every other pass runs as written. op.file is the
      // first box's force mesh; op.file1, when given, the second box's (unit
      // 10 follows the box whose force framebuffer is bound).
      const INJ_FS = ['#version 300 es', 'precision highp float;', 'precision highp sampler2D;',
                      'uniform sampler2D uInject;', 'out vec4 o;',
                      'void main() { o = texelFetch(uInject, ivec2(gl_FragCoord.xy), 0); }'].join('\n');
      NB.setProg('cell', CC.glsl.QUAD_VS, INJ_FS);
      const cp = NB.progs.cell;
      gl.useProgram(cp.p); gl.uniform1i(cp.u.uInject, 10);
      const tex = (file) => {
        const t = gl.createTexture();
        gl.activeTexture(gl.TEXTURE10);
        gl.bindTexture(gl.TEXTURE_2D, t);
        gl.texParameteri(gl.TEXTURE_2D, gl.TEXTURE_MIN_FILTER, gl.NEAREST);
        gl.texParameteri(gl.TEXTURE_2D, gl.TEXTURE_MAG_FILTER, gl.NEAREST);
        const data = file ? readF32(path.dirname(file), path.basename(file)) : new Float32Array(512 * 513 * 4);
        gl.texImage2D(gl.TEXTURE_2D, 0, gl.RGBA32F, 512, 513, 0, gl.RGBA, gl.FLOAT, data);
        return t;
      };
      injTex = tex(op.file);
      if (op.file1) {
        const t1 = tex(op.file1);
        injectFor = new Map([[NB.pm[0].forceFbo.__h, injTex], [NB.pm[1].forceFbo.__h, t1]]);
        gl.bindTexture(gl.TEXTURE_2D, injTex);
      }
      gl.activeTexture(gl.TEXTURE0);
      break;
    }
    case 'mark': trace.push({ op: '__mark', args: [], tag: op.tag }); break;
    case 'upload_tex': {   // overwrite one engine texture (e.g. the second box's force mesh)
      const t = texOf(op.name);
      upload(t.tex, t.w, t.h, readF32(path.dirname(op.file), path.basename(op.file)), gl.RGBA);
      break;
    }
    case 'dump': {
      trace.push({ op: '__dump', args: [], tag: op.tag,
                   dump: op.what.map((nm) => { const t = texOf(nm); return { name: nm, tex: t.tex.__h, w: t.w, h: t.h }; }) });
      break;
    }
    // the crater page: build (its S fields from op.opt), settle, approach, step
    case 'crater_build': {
      crater = craterPage();
      Object.assign(crater.S, op.opt);
      crater.build();
      const s = crater.sim;
      const rec = {};
      for (const k of Object.keys(s)) {
        const v = s[k];
        if (typeof v === 'number' || typeof v === 'string' || Array.isArray(v)) rec[k] = JSON.parse(JSON.stringify(v));
      }
      rec.G = crater.grid.G;
      host.ops.push({ op: 'crater_build', sim: rec, consts: crater.consts });
      fs.writeFileSync(path.join(outDir, (op.tag || 'crater') + '.mat.f32'), Buffer.from(s.matR.buffer));
      break;
    }
    case 'crater_settle': crater.settleSome(op.n); break;
    case 'crater_step': crater.step(!!op.settle); break;
    case 'crater_set_sim': {
      for (const k of Object.keys(op.fields)) crater.sim[k] = op.fields[k];
      break;
    }
    case 'crater_approach': {
      // frames until the coupling pass; the state just before that pass is dumped
      const what = op.what;
      crater.hooks.beforeCouple = () => {
        trace.push({ op: '__dump', args: [], tag: op.tagBefore,
                     dump: what.map((nm) => { const t = texOf(nm); return { name: nm, tex: t.tex.__h, w: t.w, h: t.h }; }) });
        const s = crater.sim;
        host.ops.push({ op: 'crater_couple', tag: op.tagBefore,
                        sim: { cen2: s.cen2.slice(), dir: s.dir, dImp: s.dImp, vImp: s.vImp, aImp: s.aImp, Dtc: s.Dtc, tK: s.tK } });
      };
      let frames = 0;
      while (crater.sim.phase !== 'full' && frames < 10 * crater.consts.APPROACH_FRAMES) { crater.approachStep(); frames++; }
      if (crater.sim.phase !== 'full') throw new Error('no touchdown');
      host.ops.push({ op: 'crater_touchdown', frames: frames, centre: crater.sim.centre });
      break;
    }
    default: throw new Error('unknown job op ' + op.op);
  }
}
// the pair table the pairwise pass subtracts (uploaded by init as a 64x1 RG32F texture)
{
  let tabRec = null;
  for (const r of trace) if (r.op === 'texImage2D' && r.args[3] === 64 && r.args[4] === 1 && r.args[2] === ENUM.RG32F) tabRec = r;
  host.ppTab = tabRec ? tabRec.args[8] : null;
}
host.consts = { T_UNIT: NB.T_UNIT, V_UNIT: NB.V_UNIT, E_KG: NB.E_KG, CP_ROCK: NB.CP_ROCK, CP_IRON: NB.CP_IRON, L_EM: NB.L_EM, JAM: NB.JAM, PP_CAP: NB.PP_CAP };
fs.writeFileSync(path.join(outDir, 'trace.json'), JSON.stringify({ shaders: Object.fromEntries(Array.from(shaderSrc.entries())), trace: trace }));
fs.writeFileSync(path.join(outDir, 'blobs.bin'), Buffer.concat(blobs));
fs.writeFileSync(path.join(outDir, 'host.json'), JSON.stringify(host, null, 1));
console.log('trace ops', trace.length, 'blob bytes', blobBytes);
