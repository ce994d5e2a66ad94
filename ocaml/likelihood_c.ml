(** Likelihood_c on the GPU: Felsenstein pruning over MlModel behind
    NodeData.S (lib/nodeData.ml:3-35).  Every function is TODO in the
    reference (lib/likelihood_c.ml:1-33); this module defines them.

    A value holds a node's conditional likelihoods in device memory --
    partials L[r][i][p] (FP64 or FP32) and an int32 power-of-two exponent per
    (rate class, pattern) -- plus the length of the branch from its node to
    the point where the next operation joins it (NodeData.S has no
    branch-length argument), its pattern weights and one character code per
    pattern.  Costs are -lnL (NodeData costs are minimised).  The device
    model (pi, eigen-decomposition, rates) is built once from an MlModel.t
    and shared by the values made from it. *)

open Internal

type dbuf   (** device buffer (custom block) *)
type dmodel (** device model (custom block) *)
type topo   (** compiled tree traversal (custom block) *)

type t = {
  part : dbuf;
  exps : dbuf;
  blen : float;
  n_pat : int;
  pitch : int;
  weights : dbuf;
  f32 : bool;
  mutable lnl : float;  (** lnL with this value as the root; nan until evaluated *)
  codes : int array;
  model : dmodel;
}

type m = MlModel.t
type spec = unit

external register : unit -> unit = "phylo_hip_CAML_register"
let () = register ()

external model_ : MlModel.t -> dmodel = "likelihood_CAML_model"
external weights_ : float array -> dbuf = "nonadditive_CAML_weights"
external of_tips_ : dmodel -> int array -> bool -> dbuf * dbuf * int = "likelihood_CAML_of_tips"
external median_2_ : t option -> t -> t -> dbuf * dbuf = "likelihood_CAML_median2"
external median_3_ : t -> t -> t -> dbuf * dbuf * float = "likelihood_CAML_median3"
external root_ : t -> float = "likelihood_CAML_root"
external edge_ : t -> t -> float -> float array = "likelihood_CAML_edge"
external adjust_3_ : t -> t -> float -> float array = "likelihood_CAML_adjust3"

(** The device model of an MlModel.t (build once, share across values). *)
let device_model (m : MlModel.t) = model_ m

(** median_1 of a leaf: its tip's state set per pattern (bit i = state i;
    all bits = missing data), its branch length and the pattern weights. *)
let of_tips ?codes ?(f32 = false) dm (tips : int array) (blen : float) (weights : float array) =
  let n = Array.length tips in
  if Array.length weights <> n then invalid_arg "of_tips: one weight per pattern";
  let part, exps, pitch = of_tips_ dm tips f32 in
  let codes = match codes with Some c -> c | None -> Array.init n (fun i -> i) in
  { part; exps; blen; n_pat = n; pitch; weights = weights_ weights; f32; lnl = nan; codes;
    model = dm }

let with_blen t blen = { t with blen; lnl = t.lnl }

(* -- NodeData.S: character bookkeeping ------------------------------------- *)
let cardinal t = t.n_pat
let get_codes t = Array.fold_left (fun s c -> IntSet.add c s) IntSet.empty t.codes
let mem codes t = match codes with
  | None -> true
  | Some cs -> List.exists (fun c -> Array.exists (( = ) c) t.codes) cs
let recode f t = { t with codes = Array.map f t.codes }
let filter_codes _ _ = failwith "Likelihood_c.filter_codes: gather the patterns on the host"
let filter_codes_comp _ _ = failwith "Likelihood_c.filter_codes_comp: gather the patterns on the host"
let compare a b =
  match Stdlib.compare a.codes b.codes with
  | 0 -> Stdlib.compare a.blen b.blen
  | c -> c
let union _ _ _ = failwith "TODO"

(* -- NodeData.S: medians, costs, distances --------------------------------- *)
let median_1 _ _ x = x

(** median_2: (P(t_x) L_x) * (P(t_y) L_y) per rate class; the result keeps
    the previous value's branch length (0 for a new node: set it with
    [with_blen] once the tree's length is known). *)
let median_2 _ prev x y =
  let part, exps = median_2_ prev x y in
  let blen = match prev with Some p -> p.blen | None -> 0.0 in
  { x with part; exps; blen; lnl = nan }

(** median_3: the unrooted root of three neighbours, with the tree's lnL. *)
let median_3 _ prev a b c =
  let part, exps, lnl = median_3_ a b c in
  let blen = match prev with Some p -> p.blen | None -> 0.0 in
  { a with part; exps; blen; lnl }

let median_n _ _ _ _ = failwith "TODO"

let root_cost t =
  if Float.is_nan t.lnl then t.lnl <- root_ t;
  -. t.lnl

let cost = root_cost
let leaf_cost _ = 0.0

(** distance_1: -lnL of a and b joined across their branches (one edge of
    length t_a + t_b). *)
let distance_1 _ a b = -. (edge_ a b (a.blen +. b.blen)).(0)

(** distance_2: -lnL of a, b and c joined at one point (c inserted on the
    edge between a and b). *)
let distance_2 _ a b c = let _, _, lnl = median_3_ a b c in -. lnl

(** adjust_3 m codes n a b c (as Node.Make1D.readjust_3 calls it,
    lib/node.ml:239-256): n's value is recomputed from its children a and b,
    then the length of the edge between n and c -- the value of the rest of
    the tree at n's parent -- is optimised by Newton-Raphson.  The codes of
    n are reported changed when the length moved. *)
let adjust_3 m codes n a b c =
  let touched = match codes with
    | None -> true
    | Some cs -> Array.exists (fun x -> IntSet.mem x cs) n.codes in
  if not touched then n, IntSet.empty
  else begin
    let d = { (median_2 m None a b) with blen = n.blen } in
    let r = adjust_3_ d c n.blen in
    let t' = r.(0) in
    let out = { d with blen = t'; lnl = nan } in
    let ch = if t' = n.blen then IntSet.empty
      else match codes with
        | None -> get_codes n
        | Some cs -> IntSet.inter cs (get_codes n) in
    out, ch
  end

let adjust_n _ _ _ _ = failwith "TODO"

let to_string t =
  Printf.sprintf "Likelihood_c(%d patterns, %s, blen %g)" t.n_pat
    (if t.f32 then "FP32" else "FP64") t.blen

(* -- whole-tree evaluation --------------------------------------------------- *)
type tree_data = { tips : dbuf; stride : int; n_patterns : int; tree_weights : dbuf }

external topo_create : int -> int array -> int -> topo = "phylo_hip_CAML_topo_create"
external tree_rows_ : int array array -> int -> dbuf * int = "nonadditive_CAML_tree_rows"
external derivs_ : topo -> dmodel -> tree_data -> float array -> bool -> float * float array * float array
  = "likelihood_CAML_derivs"

(** lnL of a whole tree with dlnL/dt and d2lnL/dt2 of every branch (the two
    root branches share their unrooted edge's value), in one traversal:
    [tips] has one state set per pattern for each of the N leaves, [blen]
    2N-1 lengths (the root's ignored). *)
let tree_derivs ?(f32 = false) dm ~children ~root (tips : int array array) weights blen =
  let n = Array.length tips in
  let p = if n = 0 then 0 else Array.length tips.(0) in
  let tp = topo_create n children root in
  let rows, stride = tree_rows_ tips n in
  let data = { tips = rows; stride; n_patterns = p; tree_weights = weights_ weights } in
  derivs_ tp dm data blen f32
