// Batch operations of the step loop used by the serving loop (fs_server.cpp).
// Internal to libfarrsight_hip.
#pragma once

#include <cstdint>
#include <vector>

#include "fs_internal.h"

namespace fs {
namespace sim {

struct Deck;

// A deck's meshes built on the host (no device work); safe on any thread.
int build_deck_host(const fs_deck* deck, Deck** out);
void free_deck(Deck* d);
int64_t deck_points(const Deck* d);
// Estimated device bytes the deck holds while resident in a batch.
size_t deck_footprint(const Deck* d, int precision);

// build_deck_host, then each species' points and tree replaced by the files
// <sim_dir>/simulation_output/<names[k]>/... of step `iter` (fs_io.cpp): the
// deck resumes from a state written after a remesh. Host only.
int build_deck_restart_host(const fs_deck* deck, const char* sim_dir,
                            const char* const* names, int iter, Deck** out);

// An empty batch (remesh every step) that decks join and leave.
int create_empty(fs_ctx* ctx, int precision, fs_sim** out);
// Adds the decks (ownership passes to the batch) with one relayout and
// evaluates their initial field. Deck indices of the new decks are appended.
int attach_decks(fs_sim* S, std::vector<Deck*>& decks, std::vector<int>* indices);

struct SpeciesState {
    std::vector<double> xs, ps, fs, qw, es;
    std::vector<int32_t> rec;  // panel records of the final mesh
    int iter = 0;              // steps the deck had taken
    double diag[4];
};
// Copies the final state of the given decks (every species) to the host and
// removes them with one relayout. out[k] holds deck idx[k]'s species.
int detach_decks(fs_sim* S, const std::vector<int>& idx,
                 std::vector<std::vector<SpeciesState>>& out);

}  // namespace sim
}  // namespace fs
