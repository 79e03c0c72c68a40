// Type declarations of the per-q motility fit module (motility-fit.mjs).
import { Radial } from './ddm-engine';

export declare const MODEL: Readonly<{ BROWNIAN: 0; SWIMMER_3D: 1 }>;
export declare const SWIM_MIN_UM_S: number;

export interface PreparedBin { b: number; q: number; qn: Float64Array; d: Float64Array; w: number }
export interface Prepared { bins: PreparedBin[]; lags: number[]; tau: Float64Array; dt: number; L: number; tauMax: number }

export declare function lagGrid(L: number): number[];
export declare function prepare(radial: Radial, dxUm: number, fps: number): Prepared;
export declare function bounds(pr: Prepared): [number, number, number, number];

export interface BinFit { b: number; q: number; A: number; B: number; alpha: number; vbar: number; Z: number; D: number; chi2: number; iters: number }
export declare function analyzeMotilityLM(radial: Radial, opt?: {
  dxUm?: number; fps?: number; model?: number; nstarts?: number; maxIter?: number;
  /** 6 per usable bin: A, B, alpha, ln vbar, ln Z, ln D */
  seeds?: Float64Array; nnlsAmplitudes?: boolean;
}): { pr: Prepared; fits: BinFit[]; kernelMs: number };

export declare const GATES: Readonly<{ FLOW_SIGMA: number; FLOW_MIN_UM_S: number; DRIFT_ISF_MAX: number }>;
export declare function isf3d(x: number, Z: number): number;
export declare function isf2d(x: number, Z: number): number;
export interface DriftEstimate {
  Vx: number; Vy: number; mag: number; sigma: number; angleDeg: number; significant: boolean; pairs: number; samples: number;
}
/** Uniform drift from Radial.drift (readRadial with drift on); swimmer turns on the ISF lag gating. */
export declare function estimateDrift(radial: Radial, pr: Prepared, opt?: { dxUm?: number; geometry?: '3d' | '2d' },
  swimmer?: { vbar: number; Z: number } | null): DriftEstimate;
