// Drives courant_napi.node through the wasm-shaped exports that
// web/src/solver/courant.ts reads (memory.buffer, malloc, free, cem_* with byte
// offsets for pointers), the way that wrapper uses them: inputs copied in through
// malloc + Float64Array views, results read through views at returned offsets.
//
//   node harness.js exports <addon>   export list, strings and a finalized plan (no GPU)
//   node harness.js line <addon>      a small lumped-port line stepped on the GPU
//   node harness.js workers <addon>   the export check in two worker threads while the
//                                     main thread holds a plan and data in its own arena
//   node harness.js board <addon> <trace.json>
//                                     the command-line board run (web/tools/run-board.ts:53-85)
//                                     on a mode-port board: each excitation's set-up calls
//                                     replayed from the trace, then stepped in chunks of 500
//                                     to the energy-decay stop, the two ports' waves, the
//                                     2 x 2 S matrix, its 50-ohm renormalisation per frequency
//                                     and the Touchstone text
//
// Prints one JSON object.
'use strict';

const { Worker, isMainThread, parentPort, workerData } = require('worker_threads');
const [mode, addonPath] = isMainThread ? process.argv.slice(2) : [workerData.mode, workerData.addon];
const ex = require(addonPath);

function allocF64(values) {
  const p = ex.malloc(values.length * 8);
  if (!p) throw new Error('arena out of memory');
  new Float64Array(ex.memory.buffer, p, values.length).set(values);
  return p;
}
function readF64(p, n) { return Array.from(new Float64Array(ex.memory.buffer, p, n)); }
function readString(p) {
  const mem = new Uint8Array(ex.memory.buffer);
  let end = p;
  while (mem[end] !== 0) end++;
  return Buffer.from(mem.slice(p, end)).toString('utf8');
}
function lastError() { return readString(ex.cem_last_error()); }
function check(rc, what) {
  if (rc < 0) throw new Error(`${what}: ${lastError()}`);
  return rc;
}
function lines(n, h) { return Array.from({ length: n + 1 }, (_, i) => i * h); }

// grid and geometry shared with tests/test_napi.py (same calls in the same order)
function buildLine() {
  const x = lines(60, 1e-4), y = lines(40, 1e-4), z = lines(30, 5e-5);
  const px = allocF64(x), py = allocF64(y), pz = allocF64(z);
  const s = ex.cem_create(60, 40, 30, px, py, pz);
  ex.free(px); ex.free(py); ex.free(pz);
  if (!s) throw new Error(`cem_create: ${lastError()}`);
  for (let f = 0; f < 6; f++) check(ex.cem_set_boundary(s, f, 1, 6), 'set_boundary');
  check(ex.cem_material_box(s, -1, 1, -1, 1, z[6], z[12], 4.3, 0), 'material_box');
  check(ex.cem_pec_sheet(s, 2, z[6], -1, 1, -1, 1), 'pec_sheet');
  check(ex.cem_pec_sheet(s, 2, z[12], x[10], x[50], y[18], y[22]), 'pec_sheet');
  const p1 = check(ex.cem_add_lumped_port(s, 2, 10, 20, 6, 12, -1, 50), 'port');
  const p2 = check(ex.cem_add_lumped_port(s, 2, 50, 20, 6, 12, -1, 50), 'port');
  check(ex.cem_set_gaussian(s, p1, 1, 60e-12, 15e-12), 'set_gaussian');
  const f = Array.from({ length: 40 }, (_, i) => 1e9 * (i + 1));
  const pf = allocF64(f);
  check(ex.cem_set_frequencies(s, f.length, pf), 'set_frequencies');
  ex.free(pf);
  if (ex.cem_finalize(s, 0.99) !== 0) throw new Error(`finalize: ${lastError()}`);
  return { s, p1, p2 };
}

function exportsReport() {
  const out = {};
  out.first_block = ex.malloc(8);   // a fresh arena hands out its first block
  ex.free(out.first_block);
  out.names = Object.keys(ex).filter((k) => k.startsWith('cem_')).sort();
  out.malloc = typeof ex.malloc === 'function' && typeof ex.free === 'function';
  out.arena_bytes = ex.memory.buffer.byteLength;
  out.version = readString(ex.cem_version());
  const { s } = buildLine();
  out.dt = ex.cem_dt(s);
  out.nfreq = ex.cem_nfreq(s);
  const N = 61 * 41 * 31;
  out.cb_sum = readF64(ex.cem_plan_array(s, 1, 2), N).reduce((a, b) => a + b, 0);
  out.ide_y = readF64(ex.cem_plan_array(s, 2, 1), 41);
  // error path through the arena string
  out.bad_port = ex.cem_add_lumped_port(s, 2, 10, 20, 6, 12, -1, 50);
  out.error = lastError();
  // pointer arguments outside the arena's allocated blocks are refused with a RangeError
  const rejected = (fn) => { try { fn(); return false; } catch (e) { return e instanceof RangeError; } };
  out.bad_ptr = [out.arena_bytes + 64, -8, 1.5, out.arena_bytes - 64].map(
    (p) => rejected(() => ex.cem_plan_port(s, 0, p, 0, 0)));
  // allocator: blocks are distinct, freed blocks are reused
  const a = ex.malloc(1000), b = ex.malloc(1000);
  ex.free(a);
  const c = ex.malloc(500);
  out.alloc = { distinct: a !== b && a > 0 && b > 0, reused: c === a };
  ex.free(b); ex.free(c);
  ex.cem_destroy(s);
  return out;
}


// Replays recorded courant.h calls (tests/napi_trace.py): scalars as given, simulation
// handles by creation order, arrays and structs copied into arena blocks, output
// buffers allocated with their initial bytes and read back after the call.
function replay(calls) {
  const sims = [], results = [], outputs = [];
  for (const [name, args, res] of calls) {
    const blocks = [], outs = [];
    const vals = args.map((a) => {
      if ('i' in a) return a.i;
      if ('d' in a) return a.d;
      if ('sim' in a) return sims[a.sim];
      if ('null' in a) return 0;
      const raw = Buffer.from('in' in a ? a.in : a.init, 'base64');
      const p = ex.malloc(Math.max(raw.length, 8));
      if (!p) throw new Error('arena out of memory');
      new Uint8Array(ex.memory.buffer, p, raw.length).set(raw);
      blocks.push(p);
      if ('out' in a) outs.push([p, a.out]);
      return p;
    });
    const r = ex[name](...vals);
    if (name === 'cem_create') {
      if (!r && r !== 0) throw new Error(`cem_create: ${lastError()}`);
      sims.push(r);
      results.push({ sim: sims.length - 1 });
    } else {
      results.push(typeof r === 'number' ? r : null);
    }
    outputs.push(outs.map(([p, n]) => Buffer.from(new Uint8Array(ex.memory.buffer, p, n)).toString('base64')));
    blocks.forEach((p) => ex.free(p));
  }
  return { sims, results, outputs };
}

const cdiv = (x, y) => {   // run-board.ts's div
  const d = y.re * y.re + y.im * y.im;
  return { re: (x.re * y.re + x.im * y.im) / d, im: (x.im * y.re - x.re * y.im) / d };
};

function modePortWaves(s, port, nf) {   // courant.ts modePortWaves
  const pa = ex.malloc(16 * nf), pb = ex.malloc(16 * nf), pz = ex.malloc(8 * nf), pbeta = ex.malloc(8 * nf);
  check(ex.cem_mode_port_waves(s, port, pa, pb, pz, pbeta), 'mode_port_waves');
  const a = readF64(pa, 2 * nf), b = readF64(pb, 2 * nf);
  const w = {
    a: Array.from({ length: nf }, (_, i) => ({ re: a[2 * i], im: a[2 * i + 1] })),
    b: Array.from({ length: nf }, (_, i) => ({ re: b[2 * i], im: b[2 * i + 1] })),
    z0: readF64(pz, nf),
  };
  [pa, pb, pz, pbeta].forEach((p) => ex.free(p));
  return w;
}

function renormalize2(S, z1, z2, zref) {   // courant.ts renormalize2
  const p = allocF64(S.flatMap((c) => [c.re, c.im]));
  ex.cem_s2_renormalize(p, z1, z2, zref);
  const v = readF64(p, 8);
  ex.free(p);
  return [0, 1, 2, 3].map((k) => ({ re: v[2 * k], im: v[2 * k + 1] }));
}

function jsTouchstone(f, s11, s21, s12, s22, title, version, date) {   // web/src/app/touchstone.ts:7-23
  const lines = [
    `! Courant EM ${version} — ${title}`,
    `! ${date}; reference 50; both ports excited`,
    '# HZ S RI R 50',
  ];
  f.forEach((fk, i) => {
    const a = s11[i], b = s21[i], c = s12[i], d = s22[i];
    lines.push([fk, a.re, a.im, b.re, b.im, c.re, c.im, d.re, d.im].map((v, k) => (k ? v.toExponential(9) : v.toExponential(6))).join(' '));
  });
  return lines.join('\n') + '\n';
}

function runBoard(trace) {
  const waves = [], runs = [];
  let steps = 0;
  for (const run of trace.runs) {
    const rep = replay(run.calls);
    const s = rep.sims[0];
    const dt = ex.cem_dt(s), nf = ex.cem_nfreq(s);
    let peak = 0;
    const energies = [];
    for (;;) {   // run-board.ts:64-72
      check(ex.cem_step(s, 500), 'step');
      const W = ex.cem_energy(s);
      energies.push(W);
      peak = Math.max(peak, W);
      const n = ex.cem_step_count(s);
      if ((n * dt > trace.tstop && W < trace.decay * peak) || n * dt > trace.tmax) break;
    }
    waves.push({ w1: modePortWaves(s, run.ports[0], nf), w2: modePortWaves(s, run.ports[1], nf) });
    steps += ex.cem_step_count(s);
    runs.push({ results: rep.results, outputs: rep.outputs, energies, steps: ex.cem_step_count(s) });
    ex.cem_destroy(s);
  }
  const { w1, w2 } = waves[0];
  const s11 = w1.b.map((b, i) => cdiv(b, w1.a[i])), s21 = w2.b.map((b, i) => cdiv(b, w1.a[i]));
  const s12 = waves[1].w1.b.map((b, i) => cdiv(b, waves[1].w2.a[i]));
  const s22 = waves[1].w2.b.map((b, i) => cdiv(b, waves[1].w2.a[i]));
  const r50 = s11.map((a, i) => renormalize2([a, s12[i], s21[i], s22[i]], w1.z0[i], w2.z0[i], 50));
  const flat = (v) => v.flatMap((c) => [c.re, c.im]);
  const version = readString(ex.cem_version());
  return {
    runs, steps, z1: w1.z0, z2: w2.z0,
    s11: flat(s11), s21: flat(s21), s12: flat(s12), s22: flat(s22),
    s50: r50.map((r) => flat(r)),
    touchstone: jsTouchstone(trace.f, r50.map((r) => r[0]), r50.map((r) => r[2]), r50.map((r) => r[1]),
      r50.map((r) => r[3]), trace.title, version, new Date().toISOString()),
  };
}

function emit(out) { process.stdout.write(JSON.stringify(out) + '\n'); }

let out = {};
if (!isMainThread) {
  parentPort.postMessage(exportsReport());
} else if (mode === 'exports') {
  out = exportsReport();
} else if (mode === 'workers') {
  // the main thread's own simulation and data, held across the workers' lifetimes
  const pattern = Array.from({ length: 4096 }, (_, i) => Math.sin(i) * 1e3);
  const held = allocF64(pattern);
  const { s } = buildLine();
  const dt = ex.cem_dt(s);
  const run = () => new Promise((resolve, reject) => {
    const w = new Worker(__filename, { workerData: { mode: 'exports', addon: addonPath } });
    w.once('message', resolve);
    w.once('error', reject);
  });
  Promise.all([run(), run()]).then((reports) => {
    const mine = exportsReport();
    delete mine.first_block;   // the main arena holds `held`, so its first block is taken
    const firsts = reports.map((r) => r.first_block);
    reports.forEach((r) => delete r.first_block);
    const same = reports.every((r) => JSON.stringify(r) === JSON.stringify(mine));
    const now = readF64(held, pattern.length);
    emit({
      workers: reports.length,
      workers_equal_main: same,
      worker_first_blocks: firsts,
      pattern_ok: now.every((v, i) => v === pattern[i]),
      plan_ok: ex.cem_dt(s) === dt && ex.cem_nfreq(s) === 40,
    });
    ex.cem_destroy(s);
    ex.free(held);
  }, (err) => { process.stderr.write(String(err.stack || err) + '\n'); process.exit(1); });
} else if (mode === 'line') {
  const { s, p1, p2 } = buildLine();
  // write-then-step through a held view of a library-owned field (courant.ts setField)
  const N = 61 * 41 * 31;
  const ez = new Float64Array(ex.memory.buffer, ex.cem_field(s, 2), N);
  ez[ex.cem_index(s, 30, 10, 9)] = 1e-3;
  for (let k = 0; k < 4; k++) ex.cem_step(s, 500);   // run-board.ts steps in chunks of 500
  out.steps = ex.cem_step_count(s);
  out.energy = ex.cem_energy(s);
  const nf = ex.cem_nfreq(s);
  out.v1 = readF64(ex.cem_port_v(s, p1), 2 * nf);
  out.i1 = readF64(ex.cem_port_i(s, p1), 2 * nf);
  out.v2 = readF64(ex.cem_port_v(s, p2), 2 * nf);
  out.i2 = readF64(ex.cem_port_i(s, p2), 2 * nf);
  // step-then-read through the same held view
  out.ez_probe = ez[ex.cem_index(s, 30, 20, 9)];
  ex.cem_destroy(s);
} else if (mode === 'board') {
  out = runBoard(JSON.parse(require('fs').readFileSync(process.argv[4], 'utf8')));
} else {
  throw new Error(`unknown mode ${mode}`);
}
if (isMainThread && mode !== 'workers') emit(out);
