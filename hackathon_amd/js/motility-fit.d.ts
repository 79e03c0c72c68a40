// Type declarations of the per-q motility fit module (motility-fit.mjs).
import { Radial } from './ddm-engine';

export declare const MODEL: Readonly<{ BROWNIAN: 0; SWIMMER_3D: 1; SWIMMER_2D: 2 }>;
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
  /** uniform drift speed (um/s) held fixed in every bin's model: E = exp(-D q^2 tau) J0(q vmag tau) */
  vmag?: number;
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

export type Verdict = 'MOTILE' | 'NOT_MOTILE' | 'DIRECTED_TRANSPORT' | 'NO_DYNAMICS' | 'UNLOCKED_EXPOSURE';
export declare const VERDICTS: readonly Verdict[];
export interface Gate { id: string; name: string; value: number; threshold: string; pass: boolean; sigma?: number; resolvedBins?: number }
export interface ModelFit { vbar: number; Z: number; Dd: number; Vmag: number; alpha: number; rss: number;
  amps: { A: number; B: number; W: number; P: number }[] }
export interface MotilityResult {
  verdict: Verdict; physicsVerdict: Verdict; gates: Gate[]; TSI: number;
  alpha?: number; alphaResolved?: boolean; resolvedBins?: number; vbar?: number; sigmaV?: number; Z?: number;
  Ddiff?: number; DdiffBrownian?: number; dBIC?: number; drift?: DriftEstimate; mu?: number;
  scaling?: { q: number[]; tauHalf: number[]; mu: number }; geometry?: '3d' | '2d';
  fit?: { swimmer: ModelFit; brownian: ModelFit }; q?: number[]; bins?: number[]; lags?: number[]; tau?: number[];
  nData?: number; nEff?: number; residualRho?: number;
}
/** The reference's analyzeMotility on a Radial record with drift sums (fits, drift, gates, verdict). */
export declare function analyzeMotility(radial: Radial, opt?: { dxUm?: number; fps?: number; geometry?: '3d' | '2d';
  exposureLocked?: boolean }): MotilityResult;
