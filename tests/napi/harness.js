// Drives courant_napi.node through the wasm-shaped exports that
// web/src/solver/courant.ts reads (memory.buffer, malloc, free, cem_* with byte
// offsets for pointers), the way that wrapper uses them: inputs copied in through
// malloc + Float64Array views, results read through views at returned offsets.
//
//   node harness.js exports <addon>   export list, strings and a finalized plan (no GPU)
//   node harness.js line <addon>      a small lumped-port line stepped on the GPU
//   node harness.js workers <addon>   the export check in two worker threads while the
//                                     main thread holds a plan and data in its own arena
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
} else {
  throw new Error(`unknown mode ${mode}`);
}
if (isMainThread && mode !== 'workers') emit(out);
