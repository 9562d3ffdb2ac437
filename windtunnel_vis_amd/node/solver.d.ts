// Types of the Node host (solver.mjs): the reference's SolverParams / SolverStats / GridDims
// (src/solver/types.ts) and the LbmSolver contract (src/backend/types.ts:7-18).
export type GroundMode = 'freeslip' | 'noslip' | 'moving';

export interface SolverParams {
  u: number;
  nu: number;
  cs: number;
  ground: GroundMode;
  sides: 'freeslip' | 'periodic';
  spongeNu: number;
  spongeStart: number;
  spongeIn?: number;
  tauWall?: number;
  precision?: 'f32' | 'f16' | 'bf16';
  interp?: boolean;
  collision?: 'regularized' | 'bgk';
  closure?: 'smagorinsky' | 'mlp';
  emaAlpha: number;
}

export interface SolverStats {
  fx: number; fy: number; fz: number;
  maxU: number;
  rhoRef: number;
  unstable: boolean;
  step: number;
}

export interface GridDims { nx: number; ny: number; nz: number; }

export interface WheelDef { c: number[]; axis: number[]; r: number; hw: number; }

export interface VoxelInfo { frontal: number; solidCells: number; min: [number, number, number]; max: [number, number, number]; }

/** Down-sampled field snapshot (src/solver/types.ts:45-55): 4 floats per point, x fastest. */
export interface FieldExport {
  dims: [number, number, number];
  stride: number;
  vel: Float32Array;   // (ux, uy, uz, rho)
  mean: Float32Array;  // time-averaged (ux, uy, uz, rho)
  vort: Float32Array;  // (wx, wy, wz, Q)
}

export interface LbmSolver {
  readonly dims: GridDims;
  readonly n: number;
  stepCount: number;
  uploadFlags(flags: Uint32Array): void;
  uploadSdf?(d: Float32Array): void;
  stepAndSample(steps: number): Promise<SolverStats>;
  destroy(): void;
}

export type SolverFactory = (dims: GridDims, p: SolverParams) => LbmSolver;

export declare class SolverGPU implements LbmSolver {
  constructor(dims: GridDims, params: SolverParams);
  readonly dims: GridDims;
  readonly n: number;
  readonly is3D: boolean;
  readonly precision: 'f32' | 'f16' | 'bf16';
  readonly packed: boolean;
  latest: SolverStats | null;
  readonly stepCount: number;
  readonly tau0: number;
  setParams(p: Partial<SolverParams>): void;
  reset(): void;
  clearStats(): void;
  encodeSteps(steps: number): void;
  stepAndSample(steps: number): Promise<SolverStats>;
  /** Macro, reduction and (derived) derived pass; queues a statistics read (SolverGPU.ts:225-261). */
  encodePost(derived?: boolean): void;
  /** Delivers the queued reads to onStats as they complete, keeping `latest` (SolverGPU.ts:264-287). */
  flushReads(onStats?: (s: SolverStats) => void): void;
  /** device.queue.onSubmittedWorkDone() for this solver's queue (backend/webgpu.ts:120). */
  onSubmittedWorkDone(): Promise<void>;
  uploadFlags(flags: Uint32Array): void;
  readFlags(): Promise<Uint32Array>;
  uploadSdf(d: Float32Array): void;
  setWheels(list: WheelDef[]): void;
  readPopulations(): Promise<Float32Array>;
  uploadPopulations(f: Float32Array): void;
  voxelizeSdf(prims: Float32Array): VoxelInfo;
  voxelizeMesh(mesh: { positions: ArrayLike<number>; indices: ArrayLike<number> }, matrix: ArrayLike<number>, wheels?: WheelDef[]): VoxelInfo;
  derive(): void;
  exportFields(stride: number): Promise<FieldExport>;
  /** Velocity and mean records at points (x, y, z triples in cells), 4 floats per point each. */
  probe(points: ArrayLike<number>): { vel: Float32Array; mean: Float32Array };
  /** seeds x points x (x, y, z, speed or -1). */
  streamlines(rake: Rake, seeds: number, points: number, h?: number, perLine?: number, useMean?: boolean): Float32Array;
  destroy(): void;
  static bytesFor(dims: GridDims): number;
}

export declare const solverFactory: SolverFactory;

/** Seeding line a -> b; parallel lines offset by d (cells). */
export interface Rake { a: number[]; b: number[]; d?: number[]; }

export interface TracerOptions {
  count: number;
  trail?: number;
  emitter?: 'inlet' | 'rake';
  nozzles?: number;
  lines?: number;
  maxAge?: number;
  rake?: Rake;
}

export declare class Tracers {
  constructor(solver: SolverGPU, opts: TracerOptions);
  readonly count: number;
  readonly trail: number;
  update(steps: number, opts?: { advance?: boolean; fill?: boolean; seed?: number }): void;
  read(): { state: Float32Array; trail: Float32Array; head: number };
  destroy(): void;
}
