'use strict';
// Runs TopologySolver.optimize on {model, config, actions} (argv[2], JSON)
// and prints one JSON line: per-iteration compliance, final densities, and
// the outcome of pause / resume / updateConfig / cancel actions keyed by
// iteration.
const path = require('path');
const { TopologySolver } = require(path.join(__dirname, '..', '..', 'topologyoptimization_amd', 'node'));

const { model, config, actions } = JSON.parse(process.argv[2]);
const solver = new TopologySolver();
const compliance = [];
const events = [];
solver.onPaused = (msg) => {
    events.push(['paused', msg.iteration]);
    setTimeout(() => { events.push(['resume', msg.iteration]); solver.resume(); }, 20);
};
// volumetric snapshots seen: [source, iteration, maxStress, max of the stress field, length]
const volumetric = [];
const volSummary = (src, v) => {
    if (!v) { volumetric.push([src, null]); return; }
    let m = 0;
    for (let i = 0; i < v.stress.length; i++) m = Math.max(m, v.stress[i]);
    volumetric.push([src, v.iteration, v.maxStress, m, v.stress.length]);
};
solver.onVolumetric = (msg) => volSummary('volumetric', msg.volumetricData);
const acts = actions || {};
solver.optimize(model, config, (it, c, meshData, timing, maxStress, vdata) => {
    compliance.push(c);
    if (vdata) volSummary('progress', vdata);
    const a = acts[String(it)];
    if (a === 'pause') solver.pause();
    else if (a === 'cancel') solver.cancel();
    else if (a === 'volumetric') solver.requestVolumetricSnapshot();
    else if (a && typeof a === 'object') solver.updateConfig(a);
}).then((res) => {
    volSummary('complete', res.volumetricData);
    process.stdout.write(JSON.stringify({
        ok: true, compliance, events, iterations: res.iterations, finalCompliance: res.finalCompliance,
        densities: Array.from(res.densities), maxStress: res.maxStress, feaOnly: !!res.feaOnly,
        elementStress: res.elementStress ? Array.from(res.elementStress) : null, volumetric,
    }) + '\n');
}, (err) => {
    process.stdout.write(JSON.stringify({ ok: false, error: err.message, compliance, events }) + '\n');
});
