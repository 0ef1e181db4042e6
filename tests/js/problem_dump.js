'use strict';
// Prints the problem setup problem.js derives from a reference-style
// {model, config, twod} (argv[2], JSON): topo_config, fixed DOFs, loads and
// preserved / void elements (twod: the 2-D worker's setup). No GPU work.
const path = require('path');
const NODE = path.join(__dirname, '..', '..', 'topologyoptimization_amd', 'node');
const addon = require(path.join(NODE, 'topo_addon.node'));
const { buildProblem, toTopoConfig, buildProblem2d, toTopoConfig2d } = require(path.join(NODE, 'problem.js'));

const { model, config, twod } = JSON.parse(process.argv[2]);
const p = twod ? buildProblem2d(addon, model, config) : buildProblem(addon, model, config);
const cfg = twod ? toTopoConfig2d(addon, model, config) : toTopoConfig(addon, model, config);
const idx = (a) => { const o = []; for (let i = 0; i < a.length; i++) if (a[i]) o.push(i); return o; };
const F = {};
for (let i = 0; i < p.F.length; i++) if (p.F[i] !== 0) F[i] = p.F[i];
process.stdout.write(JSON.stringify({
    config: cfg, fixed: idx(p.fixedMask), F, preserved: idx(p.preserved),
    void: p.voidMask ? idx(p.voidMask) : null, x0: p.x0 ? Array.from(p.x0) : null,
    version: addon.version(),
}) + '\n');
