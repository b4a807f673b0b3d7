export { chatterAt, timeDomainLimit, timeDomainLimits, timeDomainLimitsAsync, timeDomainLimitsBracketed, CHATTER_IDX } from './stabtd.js';
export { stabilityLobes, alimAt, directionalCoeffs } from './sld.js';
export { modeSet, frf } from './frf.js';
export { Stock, tipLut, LUT_DR } from './stock.js';
export { engagementBatch, fillEngagement, immersion } from './engage.js';
export { native } from './native.js';
