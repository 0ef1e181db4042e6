'use strict';
/**
 * Node entry point of the MI355X SIMP path.
 *
 *   const { TopologySolver, createWasmFacade, addon } = require('topologyoptimization_amd/node');
 *
 * The N-API addon (topo_addon.node, built in-tree next to this file) binds
 * libtopo_hip.so; loading fails loudly when it has not been built -- there
 * is no JavaScript fallback for the compute path.
 */
const path = require('path');

let addon;
try {
    addon = require(path.join(__dirname, 'topo_addon.node'));
} catch (err) {
    throw new Error('topo_addon.node is not built (python -m topologyoptimization_amd.build): ' + err.message);
}

const { TopologySolver } = require('./topology-solver.js');
const { createWasmFacade } = require('./wasm-facade.js');

module.exports = { TopologySolver, createWasmFacade, addon, default: TopologySolver };
