(** NonAdditive_c on the GPU: Fitch parsimony over 4-bit state sets packed
    16 sites per 64-bit word in device memory, behind NodeData.S
    (lib/nodeData.ml:3-35).  Replaces lib/nonAdditive_c.ml, whose only
    implemented operation is median_2 (lib/nonAdditive_c.ml:19-35); the
    others are TODO there and follow lib/node.ml's use of them here.

    A value holds one character block: its device states, the device pattern
    weights it shares with its siblings, its site count, its node-local
    weighted cost (lib/nonAdditive_c.ml:31) and one character code per site.
    The device buffers belong to custom blocks freed by the GC
    (ocaml/phylo_hip_stubs.c). *)

open Internal

type dbuf  (** device buffer (custom block) *)
type topo  (** compiled tree traversal (custom block) *)

type t = {
  states : dbuf;
  weights : dbuf;
  n_sites : int;
  cost : float;
  codes : int array;
}

type m = PModel.t
type spec = t

external register : unit -> unit = "phylo_hip_CAML_register"
let () = register ()

external gc_freq : int -> unit = "phylo_hip_CAML_custom_max"
external of_codes_ : int array -> dbuf = "nonadditive_CAML_of_codes"
external weights_ : float array -> dbuf = "nonadditive_CAML_weights"
external to_codes_ : dbuf -> int -> int array = "nonadditive_CAML_to_codes"
external median_2_ : t option -> t -> t -> dbuf * float = "nonadditive_CAML_median2"
external median_3_ : t option -> t -> t -> t -> dbuf * float = "nonadditive_CAML_median3"
external distance_1_ : t -> t -> float = "nonadditive_CAML_distance1"
external distance_2_ : t -> t -> t -> float = "nonadditive_CAML_distance2"
external adjust_3_ : t -> t -> t -> t -> dbuf * bool = "nonadditive_CAML_adjust3"
external select_ : t -> int array -> dbuf * dbuf = "nonadditive_CAML_select"
external topo_create : int -> int array -> int -> topo = "phylo_hip_CAML_topo_create"
external tree_rows_ : int array array -> int -> dbuf * int = "nonadditive_CAML_tree_rows"
external downpass_ : topo -> dbuf -> int -> dbuf -> int -> float * int array
  = "nonadditive_CAML_downpass"
external uppass_ : topo -> dbuf -> dbuf -> int -> int -> unit = "nonadditive_CAML_uppass"
external row_ : dbuf -> int -> int -> int -> int array = "nonadditive_CAML_row"

(** A character block from one state set per site (bit i = state i, as the
    codes of Alphabet.dna, lib/alphabet.ml:301-307) and its weights. *)
let of_sets ?codes (sets : int array) (weights : float array) : t =
  let n = Array.length sets in
  if Array.length weights <> n then invalid_arg "of_sets: one weight per site";
  let codes = match codes with Some c -> c | None -> Array.init n (fun i -> i) in
  { states = of_codes_ sets; weights = weights_ weights; n_sites = n; cost = 0.0; codes }

let to_sets t = to_codes_ t.states t.n_sites

(* -- NodeData.S: character bookkeeping (lib/nodeData.ml:9-16) -------------- *)
let cardinal t = t.n_sites

let get_codes t = Array.fold_left (fun s c -> IntSet.add c s) IntSet.empty t.codes

let mem codes t = match codes with
  | None -> true
  | Some cs -> List.exists (fun c -> Array.exists (( = ) c) t.codes) cs

let recode f t = { t with codes = Array.map f t.codes }

let select keep t =
  let idx = ref [] in
  Array.iteri (fun i c -> if keep c then idx := i :: !idx) t.codes;
  match List.rev !idx with
  | [] -> None
  | l ->
    let idx = Array.of_list l in
    let states, weights = select_ t idx in
    Some { states; weights; n_sites = Array.length idx; cost = 0.0;
           codes = Array.map (fun i -> t.codes.(i)) idx }

let filter_codes cs t = select (fun c -> IntSet.mem c cs) t
let filter_codes_comp cs t = select (fun c -> not (IntSet.mem c cs)) t

let compare a b =
  match Stdlib.compare a.codes b.codes with
  | 0 -> Stdlib.compare (to_sets a) (to_sets b)
  | c -> c

let union _ _ _ = failwith "TODO"  (* TODO in the reference too; Node.Make1D never calls it *)

(* -- NodeData.S: medians and costs ----------------------------------------- *)
let median_1 _ _ x = x

let median_2 _ prev x y =
  let states, cost = median_2_ prev x y in
  { x with states; cost }

let median_3 _ prev a b c =
  let states, cost = median_3_ prev a b c in
  { a with states; cost }

let median_n _ _ _ _ = failwith "TODO"

(** adjust_3 m codes n a b c: the Fitch (1971) final sets of n from its
    children's preliminary sets a, b and its parent's final sets c; the
    codes that changed are n's (limited to [codes]) when any site did. *)
let adjust_3 _ codes n a b c =
  let states, changed = adjust_3_ n a b c in
  let ch =
    if not changed then IntSet.empty
    else match codes with
      | None -> get_codes n
      | Some cs -> IntSet.inter cs (get_codes n)
  in
  { n with states }, ch

let adjust_n _ _ _ _ = failwith "TODO"

let cost t = t.cost
let root_cost t = t.cost
let leaf_cost _ = 0.0

let distance_1 _ a b = distance_1_ a b
let distance_2 _ a b c = distance_2_ a b c

let to_string t =
  Printf.sprintf "NonAdditive_c(%d sites, cost %g)" t.n_sites t.cost

(* -- whole-tree passes ------------------------------------------------------- *)
(** Node.Make1D.median_2 at every node of a tree in one launch: [leaves] are
    the N leaves' state sets (one int per site), [children] the rooted
    layout (internal node N + k has children.(2k), children.(2k+1)).
    Returns the weighted total cost, the number of union events per site,
    and the preliminary and final sets of every node (row v is node v). *)
let diagnose ~children ~root (leaves : int array array) (weights : float array) =
  let n = Array.length leaves in
  let s = if n = 0 then 0 else Array.length leaves.(0) in
  let tp = topo_create n children root in
  let rows, stride = tree_rows_ leaves (2 * n - 1) in
  let final, _ = tree_rows_ leaves (2 * n - 1) in
  let w = weights_ weights in
  let total, site_costs = downpass_ tp rows stride w s in
  uppass_ tp rows final stride s;
  let prelim = Array.init (2 * n - 1) (fun v -> row_ rows stride v s) in
  let fin = Array.init (2 * n - 1) (fun v -> row_ final stride v s) in
  total, site_costs, prelim, fin
