/* lolahip_registry.c -- RabbitCT's algorithm registry with the LolaHIP row.
 *
 * This file is the registry-side change a RabbitCT maintainer makes to adopt
 * LolaHIP: the registry interface of algorithmRegistry.h:10-21 (three global
 * function pointers, find-by-name, list) over the reference's table of
 * algorithmRegistry.c:39-48 with one more row,
 *     { "LolaHIP", lolaHipPrepare, lolaHipBackprojection, lolaHipFinish }.
 * It replaces algorithmRegistry.c when the reference runner is linked against
 * liblolahip.so; every other reference source is compiled unchanged.
 * Compiled against the reference's own headers.
 */
#include <stdio.h>
#include <string.h>

#include "algorithmRegistry.h"
#include "lolahip.h"

extern int lolaOmpPrepare(RabbitCtGlobalData *);
extern int lolaOmpBackprojection(RabbitCtGlobalData *);
extern int lolaOmpFinish(RabbitCtGlobalData *);
extern int lolaBunnyPrepare(RabbitCtGlobalData *);
extern int lolaBunnyBackprojection(RabbitCtGlobalData *);
extern int lolaBunnyFinish(RabbitCtGlobalData *);
extern int lolaOptPrepare(RabbitCtGlobalData *);
extern int lolaOptBackprojection(RabbitCtGlobalData *);
extern int lolaOptFinish(RabbitCtGlobalData *);
extern int lolaAsmPrepare(RabbitCtGlobalData *);
extern int lolaAsmBackprojection(RabbitCtGlobalData *);
extern int lolaAsmFinish(RabbitCtGlobalData *);

FncPrepareAlgorithmType FncPrepareAlgorithm;
FncAlgorithmIterationType FncAlgorithmIteration;
FncFinishAlgorithmType FncFinishAlgorithm;

static const AlgorithmEntryType Table[] = {
  { "LolaOMP", lolaOmpPrepare, lolaOmpBackprojection, lolaOmpFinish },
  { "LolaBunny", lolaBunnyPrepare, lolaBunnyBackprojection, lolaBunnyFinish },
  { "LolaOPT", lolaOptPrepare, lolaOptBackprojection, lolaOptFinish },
  { "LolaASM", lolaAsmPrepare, lolaAsmBackprojection, lolaAsmFinish },
  { "LolaHIP", lolaHipPrepare, lolaHipBackprojection, lolaHipFinish },
};
enum { TABLE_ROWS = sizeof Table / sizeof Table[0] };

int algorithmRegistryFind(const char *name)
{
  for (int i = 0; i < TABLE_ROWS; i++) {
    if (strcmp(Table[i].name, name) != 0)
      continue;
    FncPrepareAlgorithm   = Table[i].prepare;
    FncAlgorithmIteration = Table[i].iterate;
    FncFinishAlgorithm    = Table[i].finish;
    return 1;
  }
  printf("Unknown algorithm: %s\n", name);
  algorithmRegistryList();
  return 0;
}

void algorithmRegistryList(void)
{
  printf("Available algorithms:\n");
  for (int i = 0; i < TABLE_ROWS; i++)
    printf("  %s\n", Table[i].name);
}
