'use strict';
// Renders one frame through the Node-API binding from packed blobs on disk
// (the byte layouts a TypeScript host's packers produce), then writes the
// histogram and the RGBA8 image back to the same directory.
//   node render_blobs.js <dir>
// <dir>/meta.json: {width, height, oversample, nSlots, nRows}
// <dir>/{taps,palette,xforms}.f32  <dir>/{distrib,uniforms,de,viz}.u32
const fs = require('fs');
const path = require('path');
const p3 = require('./pyr3hip.node');

function f32(file) { const b = fs.readFileSync(file); return new Float32Array(b.buffer, b.byteOffset, b.length / 4); }
function u32(file) { const b = fs.readFileSync(file); return new Uint32Array(b.buffer, b.byteOffset, b.length / 4); }

const dir = process.argv[2];
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
const r = p3.create(meta.width, meta.height, meta.oversample, f32(path.join(dir, 'taps.f32')));
p3.reset(r, f32(path.join(dir, 'palette.f32')));
p3.iterate(r, f32(path.join(dir, 'xforms.f32')), u32(path.join(dir, 'distrib.u32')),
           u32(path.join(dir, 'uniforms.u32')));
const de = fs.existsSync(path.join(dir, 'de.u32')) ? u32(path.join(dir, 'de.u32')) : null;
p3.present(r, de, u32(path.join(dir, 'viz.u32')));
fs.writeFileSync(path.join(dir, 'hist_node.u32'), Buffer.from(p3.readHistogram(r).buffer));
fs.writeFileSync(path.join(dir, 'image_node.u8'), Buffer.from(p3.readImage(r).buffer));
console.log('node render ok', meta.width + 'x' + meta.height);
