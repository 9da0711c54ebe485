/*
 * celldsp_tasklib.h -- the cellspe-tasklib task runtime, MI355X edition.
 *
 * Drop-in for cellspe-tasklib.h:33-133 (task context, return codes, the
 * celldsp_* unit-sharing and task-control calls).  An execution unit ("SPE"
 * in the reference) is a host worker thread that owns one HIP stream and its
 * device staging buffers on the GPU; a task's input/output vectors are host
 * buffers that the unit stages to HBM, processes with the gfx950 kernels and
 * writes back, exactly where the SPU DMA'd them into local store
 * (cellspe-tasklib.cpp:692-748).
 *
 * Semantics kept from the reference:
 *   - celldsp_getFreeSPE / setFreeSPE reserve and release units under one
 *     process-wide mutex; with CDSP_BEHAV_BLOCKING getFreeSPE waits for a
 *     unit instead of returning -1 (cellspe-tasklib.cpp:226-282);
 *   - celldsp_startTask posts a task (and its next_task chain) to a reserved
 *     unit and returns; celldsp_waitTask blocks and returns the last task's
 *     result code (cellspe-tasklib.cpp:288-331, 594-602);
 *   - celldsp_executeTask = reserve + start + wait + release and returns
 *     vecNoErr (0) or !vecNoErr (1) (cellspe-tasklib.cpp:338-358);
 *   - completed_flag holds CELLDSP_TASK_* and tick_count the unit's device
 *     time in CELLDSP_TIMEBASE ticks (cellspe-tasklib.cpp:583-592).
 */
#ifndef CELLDSP_TASKLIB_H
#define CELLDSP_TASKLIB_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CELLDSP_MAX_SPE_NR 8      /* execution units (cellspe-tasklib.h:24) */
#define MALLOC_QUADALIGN 7
#define ALIGNMENT_QUADALIGN 128
#define _QUAD_ALIGN __attribute__((aligned(128)))

#ifndef vecNoErr
#define vecNoErr 0
#endif

/* tick_count unit: 10 ns of device time (SPU_TIMEBASE analogue) */
#define CELLDSP_TIMEBASE 100000000
#define CELLDSP_TICKS_TO_USEC(x) ((x) * 1e6 / CELLDSP_TIMEBASE)

/* flags and return codes (cellspe-tasklib.h:37-41) */
#define CELLDSP_TASK_INCOMPLETE 0xFFFF
#define CELLDSP_TASK_OK 0x0000
#define CELLDSP_TASK_GENERICERROR 0x0001
#define CELLDSP_TASK_UNKOWNCOMMAND 0x0002

typedef union {
    unsigned long long ull;
    unsigned int ui[2];
    void* p;
} Cell_addr64;

/* task sub-contexts (cellspe-tasklib.h:52-60) */
typedef struct CellDSPParams_FFT_tt { unsigned int nfft, log2_nfft; } CellDSPParams_FFT_t;
typedef struct CellDSPParams_Phase_tt { unsigned int argcount; float phase_start, phase_delta; } CellDSPParams_Phase_t;
typedef struct CellDSPParams_Difx_tt {
    unsigned int nfft, log2_nfft, useUpperOrLower, twicenumchannels, numchannels;
} CellDSPParams_Difx_t;
/* CELLDSP_TASK_UNPACK parameters (the reference leaves the command a stub):
 * nchan interleaved channels in {1,2,4,8,16}; 0 means 1 */
typedef struct CellDSPParams_Unpack_tt { unsigned int nchan; } CellDSPParams_Unpack_t;
typedef union CellDSKSubcontext_tt {
    CellDSPParams_FFT_t FFT;
    CellDSPParams_Phase_t Phase;
    CellDSPParams_Difx_t Difx;
    CellDSPParams_Unpack_t Unpack;
} CellDSPSubcontext_t;

#define CELLDSP_MAX_NUM_IOVECTORS 8
typedef struct CellDSPTask_tt {
    unsigned int command _QUAD_ALIGN;
    volatile unsigned short completed_flag;
    unsigned int tick_count;
    unsigned int num_input_vectors;
    Cell_addr64 inputvector_ptrs[CELLDSP_MAX_NUM_IOVECTORS];
    unsigned int inputvector_lengths[CELLDSP_MAX_NUM_IOVECTORS];   /* bytes */
    unsigned int num_output_vectors;
    Cell_addr64 outputvector_ptrs[CELLDSP_MAX_NUM_IOVECTORS];
    unsigned int outputvector_lengths[CELLDSP_MAX_NUM_IOVECTORS];  /* bytes */
    CellDSPSubcontext_t subcontext;
    Cell_addr64 next_task;   /* .p = NULL if none: streamed task chain */
} CellDSPTask_t _QUAD_ALIGN;

/* unit flags (cellspe-tasklib.h:161-168) */
#define CDSP_EF_RESERVED 1
#define CDSP_EF_CONTINUABLE 2
#define CDSP_EF_DEFAULT 0
#define CDSP_BEHAV_BLOCKING 1
/* Not in the reference: a next_task chain of at least two device tasks
 * (IOTEST, MAC, MULC, ADDC, FFT_EXEC[_DYN], SINCOS*, UNPACK, DIFXPROCESS)
 * that a unit executes again with the same commands, vectors (addresses and
 * lengths) and sub-contexts is captured once into a hipGraph and replayed
 * with one launch; host vectors are mirrored through pinned staging, so data
 * passed between tasks through a host vector keeps its order.  Each task's
 * tick_count is then the chain's device time divided evenly. */
#define CDSP_BEHAV_GRAPH 2
#define CDSP_BEHAV_DEFAULT 0

/* unit sharing and task control (cellspe-tasklib.h:122-133) */
int celldsp_start(void);
int celldsp_stop(void);
int celldsp_getFreeSPE(void);
int celldsp_setFreeSPE(int targetSPE);
int celldsp_startTask(int targetSPE, CellDSPTask_t* task);
int celldsp_waitTask(int targetSPE);
int celldsp_executeTask(CellDSPTask_t* task);
int celldsp_startContinuableTask(CellDSPTask_t* task, int* speID);
int celldsp_continueTask(CellDSPTask_t* task, int* speID);
int celldsp_stopContinuableTask(int* speID);
void celldsp_setBehaviour(unsigned int flags);
unsigned int celldsp_getBehaviour(void);

/* introspection used by tests: number of free units, started client count */
int celldsp_freeCount(void);
int celldsp_clientCount(void);
/* chains executed from a captured graph since celldsp_start (CDSP_BEHAV_GRAPH) */
long celldsp_graphReplays(void);

#ifdef __cplusplus
} /* extern "C" */

/* Memory-aligned task context builder (cellspe-tasklib.h:175-222). */
class CellDSPTaskClass {
  protected:
    CellDSPTask_t task;
  public:
    explicit CellDSPTaskClass(unsigned int dspcommand)
    {
        __builtin_memset((void*)&task, 0, sizeof(task));
        task.command = dspcommand;
        task.completed_flag = CELLDSP_TASK_INCOMPLETE;
        task.next_task.p = 0;
    }
    void addInVec(void* src, size_t bytelength)
    {
        task.inputvector_ptrs[task.num_input_vectors].p = src;
        task.inputvector_lengths[task.num_input_vectors++] = (unsigned int)bytelength;
    }
    void addOutVec(void* dest, size_t bytelength)
    {
        task.outputvector_ptrs[task.num_output_vectors].p = dest;
        task.outputvector_lengths[task.num_output_vectors++] = (unsigned int)bytelength;
    }
    void setNext(CellDSPTask_t* nexttask) { task.next_task.p = (void*)nexttask; }
    void setFFTparams(unsigned int fftlen) { task.subcontext.FFT.nfft = fftlen; }
    CellDSPTask_t* getTask() { return &task; }
    CellDSPSubcontext_t* getSubcontext() { return &task.subcontext; }
};
#endif

#endif
