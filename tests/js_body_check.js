// The Node host's volume upload body (run by tests/test_level2.py): reads a header (JSON) and the codes
// (raw bytes), writes volumeBody(header, codes) to the output path.
'use strict';
const fs = require('fs');
const path = require('path');
const O = require(path.join(__dirname, '..', 'outbreaklab_amd', 'js', 'outbreak.js'));
const [hp, cp, out] = process.argv.slice(2);
const header = JSON.parse(fs.readFileSync(hp, 'utf8'));
const codes = new Uint8Array(fs.readFileSync(cp));
fs.writeFileSync(out, O.volumeBody(header, codes));
console.log(JSON.stringify({ ok: true, bytes: codes.length }));
