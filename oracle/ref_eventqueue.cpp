/*
 * ref_eventqueue.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Driver linked against the reference's own event queue,
 * Kernel/src/simulation/EventQueue.cpp and Kernel/src/simulation/Event.cpp,
 * compiled where they lie (oracle/Makefile, target `ref`).  It replays a
 * sequence of heap operations on queue 0 and writes the order in which the
 * elements leave, so that the oracle's restatement of the heap can be
 * checked against the reference's.
 *
 *   ref_eventqueue <ops.bin> <out.bin>
 *   ops.bin: int64 n, then n records {int8 op, double time, int32 priority}
 *            (op 1 inserts an element whose id is the number of earlier
 *            inserts, op 0 removes the top element)
 *   out.bin: int32 ids in removal order
 *
 * The events are this driver's own minimal Event subclass; only their time
 * and ProcessingPriority() matter to the queue.
 */
#include <cstdint>
#include <cstdio>
#include <vector>
#include "simulation/Event.h"
#include "simulation/EventQueue.h"

namespace {
class ReplayEvent : public Event {
  public:
    ReplayEvent(double t, int prio, int32_t id) : Event(t, 0), prio_(prio), id_(id) {}
    void ProcessEvent(Simulation *, RealTimeRestrictionLevel) override {}
    void ProcessEvent(Simulation *) override {}
    enum EventPriority ProcessingPriority() override { return (enum EventPriority)prio_; }
    void PrintType() override {}
    int32_t id() const { return id_; }

  private:
    int prio_;
    int32_t id_;
};
}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 3;
    int64_t n = 0;
    if (fread(&n, sizeof(n), 1, f) != 1) return 3;
    EventQueue q(1);
    std::vector<int32_t> out;
    int32_t next = 0;
    for (int64_t i = 0; i < n; i++) {
        int8_t op;
        double t;
        int32_t prio;
        if (fread(&op, 1, 1, f) != 1 || fread(&t, 8, 1, f) != 1 || fread(&prio, 4, 1, f) != 1) return 3;
        if (op) {
            q.InsertEvent(0, new ReplayEvent(t, prio, next++));
        } else if (q.Size(0) > 0) {
            ReplayEvent *e = (ReplayEvent *)q.RemoveEvent(0);
            out.push_back(e->id());
            delete e;
        }
    }
    fclose(f);
    FILE *g = fopen(argv[2], "wb");
    if (!g) return 4;
    fwrite(out.data(), sizeof(int32_t), out.size(), g);
    fclose(g);
    return 0;
}
