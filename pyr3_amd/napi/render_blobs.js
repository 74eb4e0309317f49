'use strict';
// Renders one frame through the Node-API binding from packed blobs on disk
// (the byte layouts a TypeScript host's packers produce), then writes the
// histogram and the RGBA8 image back to the same directory.
//   node render_blobs.js <dir>
// <dir>/meta.json: {width, height, oversample, capture (read the index map too)}
// <dir>/{taps,palette,xforms}.f32  <dir>/{distrib,uniforms,de,viz}.u32
const fs = require('fs');
const path = require('path');
const p3 = require('./pyr3hip.node');

function f32(file) { const b = fs.readFileSync(file); return new Float32Array(b.buffer, b.byteOffset, b.length / 4); }
function u32(file) { const b = fs.readFileSync(file); return new Uint32Array(b.buffer, b.byteOffset, b.length / 4); }

async function main(dir) {
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
  const r = p3.create(meta.width, meta.height, meta.oversample, f32(path.join(dir, 'taps.f32')));
  p3.reset(r, f32(path.join(dir, 'palette.f32')));
  p3.iterate(r, f32(path.join(dir, 'xforms.f32')), u32(path.join(dir, 'distrib.u32')),
             u32(path.join(dir, 'uniforms.u32')));
  const de = fs.existsSync(path.join(dir, 'de.u32')) ? u32(path.join(dir, 'de.u32')) : null;
  p3.present(r, de, u32(path.join(dir, 'viz.u32')));
  // the readbacks are promises (src/renderer.ts:196-209); both are queued
  // before either is awaited
  const [hist, image] = await Promise.all([p3.readHistogram(r), p3.readImage(r)]);
  fs.writeFileSync(path.join(dir, 'hist_node.u32'), Buffer.from(hist.rgba.buffer));
  fs.writeFileSync(path.join(dir, 'image_node.u8'), Buffer.from(image.buffer));
  if (meta.capture) {
    const m = await p3.readIndexMap(r);
    fs.writeFileSync(path.join(dir, 'idx_node.u32'), Buffer.from(m.idxSum.buffer));
    fs.writeFileSync(path.join(dir, 'cnt_node.u32'), Buffer.from(m.count.buffer));
  }
  console.log('node render ok', meta.width + 'x' + meta.height, hist.superW + 'x' + hist.superH);
}

main(process.argv[2]).catch((e) => { console.error(e.message); process.exit(1); });
