// Loads the N-API addon (../mcsim.node, built by machinig-sim_amd/build.py), which binds
// libmcsim.so.  There is no JavaScript fallback: a missing addon or device throws.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
let addon = null;

export function native() {
  if (addon) return addon;
  try {
    addon = require('../mcsim.node');
  } catch (e) {
    throw new Error('mcsim.node is not built (python machinig-sim_amd/build.py): ' + e.message);
  }
  return addon;
}
