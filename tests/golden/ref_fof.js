// Golden-vector generator for the host analysis (development container only).
//
// The reference's friends-of-friends analysis `fof` (src/nbody.js:646-893:
// groups, the iterated core centre, the books, the loose list and the Morton
// permutation) is a self-contained function -- the reference itself ships it
// to a worker as text. This script takes that function's text out of the
// reference file, evaluates it in a fresh vm context and runs it on the
// arrays named by the job, writing the report as JSON and the permutation and
// loose list as binary.
//   node ref_fof.js <job.json>   job: {N, m, a, k, touch, cap, sort, P, V, R, D, out}
'use strict';
const fs = require('fs');
const path = require('path');
const vm = require('vm');

const REF = process.env.REF_DIR || '/root/reference/src';
const job = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const text = fs.readFileSync(path.join(REF, 'nbody.js'), 'utf8');
const start = text.indexOf('function fof(');
if (start < 0) throw new Error('fof not found');
let depth = 0, end = -1;
for (let i = text.indexOf('{', start); i < text.length; i++) {
  if (text[i] === '{') depth++;
  else if (text[i] === '}') { depth--; if (depth === 0) { end = i + 1; break; } }
}
const ctx = vm.createContext({});
vm.runInContext(text.slice(start, end), ctx, { filename: 'nbody.js#fof' });
function f32(file) { const b = fs.readFileSync(file); return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); }
ctx.args = [job.N, job.m, job.a, f32(job.P), f32(job.V), f32(job.R), f32(job.D), job.k, job.touch, job.cap, job.sort];
const r = vm.runInContext('fof.apply(null, args)', ctx);
const out = {};
for (const k of Object.keys(r)) if (k !== 'perm' && k !== 'pp') out[k] = r[k];
out.ppLoose = r.pp.loose;
out.ppCount = r.pp.list.length;
fs.writeFileSync(job.out + '.report.json', JSON.stringify(out));
fs.writeFileSync(job.out + '.pplist.f32', Buffer.from(new Float32Array(r.pp.list).buffer));
if (r.perm) fs.writeFileSync(job.out + '.perm.f32', Buffer.from(new Float32Array(r.perm).buffer));
console.log('fof done', r.pp.list.length, r.perm ? 'sorted' : '');
