export { chatterAt, timeDomainLimit, timeDomainLimits, timeDomainLimitsAsync, timeDomainLimitsBracketed, CHATTER_IDX } from './stabtd.js';
export { stabilityLobes, alimAt, directionalCoeffs, compositeLobes } from './sld.js';
export { modeSet, frf } from './frf.js';
export { Stock, tipLut, LUT_DR } from './stock.js';
export { engagementBatch, fillEngagement, immersion } from './engage.js';
export { toolModel, isAssembly } from './tool.js';
export { modesAt, tipFactor, bearingFactor, twoDofModes, speedDependent, SOFT_DEFAULT } from './spindle.js';
export { native } from './native.js';
export { CuttingEngine, BINS, END_BINS, END_RINGS, H_REF } from './cutting.js';
