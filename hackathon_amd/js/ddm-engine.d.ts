// Type declarations of the MI355X DDM engine facade (ddm-engine.mjs).  The
// class surface is the reference's WebGPU engine (motility-ddm
// src/ddm-engine.ts:62-426) and the Radial record of src/ddm_ref.js:336-348.

export declare const DDM: Readonly<{
  N: number; T: number; L: number; NB: number; KMIN: number; KMAX: number; KNODES: number; DRIFT_STRIDE: number;
}>;

export interface Radial {
  N: number;
  NB: number;
  L: number;
  kCenter: Float64Array;
  kNodes: Float64Array;
  nModes: Uint32Array;
  /** D[b*L + tau] */
  D: Float64Array;
  P: Float64Array;
  /** NB*L*8 phase-drift sums: (b*L + tau)*8 + {Sw, Sxx, Sxy, Syy, Rx, Ry, Rpp, count} */
  drift: Float64Array;
}

export interface Bins {
  N: number;
  NB: number;
  kmin: number;
  kmax: number;
  binOffsets: Uint32Array;
  binModes: Int32Array;
  kCenter: Float64Array;
  nModes: Uint32Array;
  kNodes: Float64Array;
}

export interface EngineOptions {
  N?: number;
  T?: number;
  L?: number;
  NB?: number;
  kmin?: number;
  kmax?: number;
  /** 0: V8 Math.hypot bin rule (reference parity), 1: correctly rounded */
  hypotMode?: number;
  knodes?: number;
  /** accepted for signature compatibility; the kernels are compiled HIP */
  shaderCode?: string;
  powerPreference?: string;
  onLost?: (info: unknown) => void;
  onError?: (err: Error) => void;
}

export declare class DdmEngine {
  static create(opts?: EngineOptions): Promise<DdmEngine>;
  readonly N: number;
  readonly T: number;
  readonly L: number;
  readonly NB: number;
  readonly bins: Bins;
  readonly adapterInfo: { vendor: string; architecture: string; description: string };
  throttled: number;
  readonly pendingSubmits: number;
  readonly saturated: boolean;
  waitForCapacity(): Promise<void>;
  ingestLuma(luma: Float32Array, slot: number): boolean;
  /** the frame stack is copied and transformed off the event loop; calls on one engine run in call order */
  ingestFrames(frames: Uint8Array | Uint16Array | Float32Array, nframes: number, firstSlot?: number): Promise<boolean>;
  computeTemporal(order: Uint32Array, mask: Float32Array, opts?: { yieldBetweenChunks?: boolean }): Promise<void>;
  readRadial(opts?: { drift?: boolean }): Promise<Radial>;
  readHistory(): Promise<Float32Array>;
  /** mean luma of the last frame ingested (u8/255, u16/65535, f32 as given) */
  readMean(): Promise<number>;
  /** per-mode D[m*L+tau], C[(m*L+tau)*2+{re,im}], P[m] for every mode (computed on demand) */
  readTemporalFull(): Promise<{ D: Float32Array; C: Float32Array; P: Float32Array }>;
  idle(): Promise<void>;
  destroy(): void;
}

export declare function linearOrder(valid: ArrayLike<boolean | number>): { order: Uint32Array; mask: Float32Array };
export declare function shuffleOrder(o: { order: Uint32Array; mask: Float32Array }, rand: () => number): { order: Uint32Array; mask: Float32Array };
export declare function mulberry32(seed: number): () => number;

export type TimeSource = 'captureTime' | 'mediaTime' | 'watchdog';
export interface Snapshot { order: Uint32Array; mask: Float32Array; used: number; newestN: number }
/** Live frame -> history-slot clock (the reference's FrameClock, ring length T). */
export declare class FrameClock {
  constructor(fps: number, T?: number);
  readonly T: number;
  readonly dtMs: number;
  readonly frames: number;
  readonly missed: number;
  readonly duplicates: number;
  readonly resets: number;
  readonly timeSource: TimeSource | null;
  readonly newestN: number;
  readonly medianDtMs: number;
  readonly validCount: number;
  reset(): void;
  place(tMs: number, source?: TimeSource): { slot: number; n: number } | null;
  drop(n: number): void;
  snapshot(opts?: { guard?: number; shuffle?: boolean; rand?: () => number }): Snapshot;
}
